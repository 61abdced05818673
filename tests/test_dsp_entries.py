"""Per-family DSP entries of the C ABI (rav2d_gpu_deblock_sb, _cdef_dir,
_cdef_filter, _ccso_prep/_add, _lr_wiener, _gdf_prep/_add,
_generate_grain_y/_uv, _fgy/_fguv_32x32xn) against the oracle's restatement
of the same reference function, bit-exact, on seeded random units at 8 and
10 bits: the function-table entries of crates/rav2d/src/dsp.rs:9-14, 86-109."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O

P = C.c_void_p


def _olib():
    lib = O.load()
    if not getattr(lib, "_dsp", False):
        sigs = {
            "oracle_deblock_sb64_16": [P, C.c_ssize_t, C.c_int, P, P, P, P, C.c_int, C.c_int, C.c_int],
            "oracle_cdef_filter_block_16": [P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_int, C.c_int],
            "oracle_ccso_prep_16": [P, C.c_int, P, C.c_ssize_t, P, P, P, C.c_uint32, C.c_int, C.c_int, C.c_uint32,
                                    C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
            "oracle_ccso_add_16": [P, C.c_ssize_t, P, C.c_int, P, P, C.c_int, C.c_int, P, C.c_int],
            "oracle_lr_ns_wiener_y": [P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, P, C.c_int, P, C.c_int],
            "oracle_lr_wiener_multi": [P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, P, P, P, P, C.c_int, C.c_int, P,
                                       C.c_int],
            "oracle_lr_ns_wiener_uv": [P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, P, P, C.c_ssize_t, P, P, C.c_int,
                                       C.c_int, C.c_int, C.c_int, P, C.c_int],
            "oracle_lr_gdf_prep": [P, C.c_int, P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int],
            "oracle_lr_gdf_add": [P, C.c_ssize_t, P, C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_int],
            "oracle_fgy_16": [P, P, C.c_ssize_t, C.POINTER(R.FilmGrainData), C.c_uint32, C.c_int, P, P, C.c_int,
                              C.c_int, C.c_int],
            "oracle_fguv_16": [P, P, C.c_ssize_t, C.POINTER(R.FilmGrainData), C.c_uint32, C.c_int, P, P, C.c_int,
                               C.c_int, P, C.c_ssize_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
        }
        for name, args in sigs.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = None
        lib._dsp = True
    return lib


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


def _dt(bd):
    return np.uint16 if bd > 8 else np.uint8


def _p(a):
    return a.ctypes.data_as(P)


def _addr16(a, y, x):
    return a.ctypes.data + 2 * (y * a.shape[1] + x)


def _gpu_and_host(ctx, img, bd):
    """Device copy of img in the bit depth's sample type, and a 16-bit host copy."""
    return ctx.device_array(img.astype(_dt(bd))), np.ascontiguousarray(img, dtype=np.uint16).copy()


def _blocky(rng, shape, bd, step=24):
    hi = (1 << bd) - 1
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    base = ((yy // 8) * 7 + (xx // 8) * 5) % 9 * (step << (bd - 8))
    return (base + (400 << (bd - 8) >> 2) + rng.integers(-(6 << (bd - 8)), 6 << (bd - 8), size=shape)).clip(0, hi)


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("dir_,uv", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_deblock_sb(ctx, bd, dir_, uv):
    rng = np.random.default_rng(10 * bd + 2 * dir_ + uv)
    lib = _olib()
    for it in range(6):
        shape = (72, 32) if dir_ == 0 else (32, 72)
        org = (4, 16) if dir_ == 0 else (16, 4)
        img = _blocky(rng, shape, bd)
        dev, host = _gpu_and_host(ctx, img, bd)
        nm = 3 if uv else 4
        vmask = np.zeros(4, dtype=np.uint16)
        for i in range(nm):
            vmask[i] = rng.integers(0, 1 << 16)
        ll = rng.integers(0, 1 << 16, size=2).astype(np.uint16) & rng.integers(0, 1 << 16, size=2).astype(np.uint16)
        q = rng.integers(1, 60, size=16).astype(np.uint16)
        side = rng.integers(1, 400, size=16).astype(np.uint16)
        edge = int(rng.integers(0, 2))
        lib.oracle_deblock_sb64_16(_addr16(host, *org), shape[1], (1 << bd) - 1, _p(vmask), _p(ll), _p(q),
                                   _p(side), edge, 1 if dir_ == 0 else 0, uv)
        R._check(ctx.lib.rav2d_gpu_deblock_sb(ctx.h, dir_, uv, dev.at(*org), shape[1] * dev.dtype.itemsize,
                                              _p(vmask), _p(ll), _p(q), _p(side), edge, bd, None), "deblock_sb")
        got = dev.read()
        dev.close()
        assert np.array_equal(got.astype(np.uint16), host), (it, dir_, uv)


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_cdef_dir(ctx, bd):
    rng = np.random.default_rng(50 + bd)
    L = O.load()
    out = ctx.device_array(np.zeros(2, dtype=np.int32))
    for it in range(40):
        yy, xx = np.mgrid[0:12, 0:12]
        ang = rng.uniform(0, np.pi)
        img = (512 + 200 * np.sin((yy * np.cos(ang) + xx * np.sin(ang)) * rng.uniform(0.3, 1.2)) +
               rng.integers(-20, 20, size=yy.shape)).clip(0, 1023)
        if bd == 8:
            img = img // 4
        dev, host = _gpu_and_host(ctx, img, bd)
        var = C.c_uint32()
        want = L.oracle_cdef_find_dir(C.c_void_p(_addr16(host, 2, 2)), 1, 12, bd - 8, C.byref(var))
        R._check(ctx.lib.rav2d_gpu_cdef_dir(ctx.h, dev.at(2, 2), 12 * dev.dtype.itemsize, bd, out.at(0),
                                            out.at(1), None), "cdef_dir")
        got = out.read()
        dev.close()
        assert (int(got[0]), int(np.uint32(got[1]))) == (want, var.value), it
    out.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("size", [(8, 8), (4, 8), (4, 4)])
def test_cdef_filter(ctx, bd, size):
    w, h = size
    rng = np.random.default_rng(70 + bd + w * 3 + h)
    lib = _olib()
    cs = bd - 8
    for it in range(24):
        edges = int(rng.integers(0, 16)) if it % 3 else 15
        W = w + 8
        img = _blocky(rng, (h + 4, W), bd, step=6)
        left = rng.integers(0, 1 << bd, size=(h, 2))
        pri = int(rng.integers(0, 16)) << cs
        sec = [0, 1, 2, 4][int(rng.integers(0, 4))] << cs
        if pri == 0 and sec == 0:
            sec = 2 << cs
        dirn = int(rng.integers(0, 8))
        damping = int(rng.integers(3, 7)) + cs - (1 if w == 4 else 0)
        dev, host = _gpu_and_host(ctx, img, bd)
        dl, hl = _gpu_and_host(ctx, left, bd)
        lib.oracle_cdef_filter_block_16(_addr16(host, 2, 2), W, _p(hl), _addr16(host, 0, 0), _addr16(host, h + 2, 0),
                                        pri, sec, dirn, damping, w, h, edges, cs)
        R._check(ctx.lib.rav2d_gpu_cdef_filter(ctx.h, dev.at(2, 2), W * dev.dtype.itemsize, dl.ptr, dev.at(0, 2),
                                               dev.at(h + 2, 2), pri, sec, dirn, damping, w, h, edges, bd, None),
                 "cdef_filter")
        got = dev.read()
        dev.close()
        dl.close()
        assert np.array_equal(got.astype(np.uint16), host), (it, edges, pri, sec, dirn)


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_ccso_prep_and_add(ctx, bd):
    rng = np.random.default_rng(90 + bd)
    lib = _olib()
    for it in range(16):
        ss = int(rng.integers(0, 2))
        w, h = [(64, 8), (32, 8), (16, 4), (48, 8)][it % 4]
        w >>= ss  # one 64-luma-sample superblock column at most
        pw, ph = w << ss, h << ss
        W = pw + 8
        luma = _blocky(rng, (ph + 4, W), bd, step=9)
        left = rng.integers(0, 1 << bd, size=(ph, 2))
        edges = int(rng.integers(0, 16))
        bo = int(rng.integers(0, 2))
        mbl = int(rng.integers(0, 8 if bo else 4))
        ext = int(rng.integers(0, 7))
        qs = int(rng.integers(1, 40)) << (bd - 8)
        eclf = int(rng.integers(0, 2))
        dev, host = _gpu_and_host(ctx, luma, bd)
        dl, hl = _gpu_and_host(ctx, left, bd)
        want = np.zeros((h, w), dtype=np.uint8)
        lib.oracle_ccso_prep_16(_p(want), w, _addr16(host, 2, 4), W, _p(hl), _addr16(host, 0, 4),
                                _addr16(host, ph + 2, 4), mbl, ext, qs, eclf, bo, edges, w, h, ss, ss, bd)
        idx = ctx.device_array(np.zeros((h, w), dtype=np.uint8))
        it_ = dev.dtype.itemsize
        R._check(ctx.lib.rav2d_gpu_ccso_prep(ctx.h, idx.ptr, w, dev.at(2, 4), W * it_, dl.ptr, dev.at(0, 4),
                                             dev.at(ph + 2, 4), mbl, ext, qs, eclf, bo, edges, w, h, ss, ss, bd,
                                             None), "ccso_prep")
        got_idx = idx.read()
        assert np.array_equal(got_idx, want), ("prep", it)
        # ccso_add on a chroma block with that index map
        plane = rng.integers(0, 1 << bd, size=(h, w))
        dp, hp = _gpu_and_host(ctx, plane, bd)
        offs = rng.integers(0, 256, size=64).astype(np.uint8)
        lut = rng.integers(-10, 11, size=8).astype(np.int8)
        ll = np.zeros((h // 4 + 1, 4), dtype=np.uint16)
        ll[:, 0] = rng.integers(0, 1 << 16, size=h // 4 + 1) & rng.integers(0, 1 << 16, size=h // 4 + 1)
        lib.oracle_ccso_add_16(_p(hp), w, _p(want), w, _p(offs), _p(lut), w, h, _p(ll), (1 << bd) - 1)
        R._check(ctx.lib.rav2d_gpu_ccso_add(ctx.h, dp.ptr, w * it_, idx.ptr, w, _p(offs), _p(lut), w, h, _p(ll), bd,
                                            None), "ccso_add")
        assert np.array_equal(dp.read().astype(np.uint16), hp), ("add", it)
        for d in (dev, dl, idx, dp):
            d.close()


def _stripe(rng, w, h, bd, pad=8):
    """Rows: 2 above, the h stripe rows, 2 below, in one image (same stride);
    `pad` columns each side; plus a separate left buffer (h rows of 6)."""
    img = _blocky(rng, (h + 4, w + 2 * pad), bd, step=5)
    left = rng.integers(0, 1 << bd, size=(h, 6))
    return img, left


def _ll_rows(rng, h, w):
    ll = np.zeros((h // 4 + 1, 4), dtype=np.uint16)
    ll[:, 0] = rng.integers(0, 1 << 16, size=h // 4 + 1) & rng.integers(0, 1 << 16, size=h // 4 + 1) & \
        rng.integers(0, 1 << 16, size=h // 4 + 1)
    return ll


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_lr_wiener(ctx, bd, kind):
    rng = np.random.default_rng(200 + bd * 4 + kind)
    lib = _olib()
    for it in range(8):
        w, h = [(64, 64), (64, 56), (40, 16), (20, 8), (32, 32), (64, 60), (12, 8), (64, 8)][it]
        if kind == 1:
            w, h = min(w, 32), min(h, 32)
        edges = int(rng.integers(0, 16)) if it else 15
        img, left = _stripe(rng, w, h, bd)
        W = img.shape[1]
        dev, host = _gpu_and_host(ctx, img, bd)
        dl, hl = _gpu_and_host(ctx, left, bd)
        ll = _ll_rows(rng, h, w)
        par = R.WienerParams()
        keep = []
        o_args = [_addr16(host, 2, 8), W, _p(hl), _addr16(host, 0, 8), _addr16(host, h + 2, 8), w, h]
        it_ = dev.dtype.itemsize
        if kind == 0:
            filt = rng.integers(-128, 128, size=16).astype(np.int8)
            par.single.filter = filt.ctypes.data
            lib.oracle_lr_ns_wiener_y(*o_args, _p(filt), edges, _p(ll), bd)
        elif kind == 1:
            filt = rng.integers(-128, 128, size=18).astype(np.int8)
            ds = it % 3
            luma = _blocky(rng, (2 * h + 4, 2 * w + 16), bd, step=7)
            dlum, hlum = _gpu_and_host(ctx, luma, bd)
            keep.append(dlum)
            LW = luma.shape[1]
            lib.oracle_lr_ns_wiener_uv(*o_args, _p(filt), _addr16(hlum, 2, 8), LW, _addr16(hlum, 0, 8),
                                       _addr16(hlum, 2 * h + 2, 8), 1, 1, ds, edges, _p(ll), bd)
            par.single.filter = filt.ctypes.data
            par.single.luma = dlum.at(2, 8)
            par.single.luma_top = dlum.at(0, 8)
            par.single.luma_bottom = dlum.at(2 * h + 2, 8)
            par.single.stride = LW * it_
            par.single.ss_ver = par.single.ss_hor = 1
            par.single.ds_flt = ds
        else:
            sub = rng.integers(0, 16 if kind == 2 else 64, size=256).astype(np.uint8)
            noskip = rng.integers(0, 1 << 16, size=20).astype(np.uint16)
            base_q = int(rng.integers(4, 1500))
            if kind == 2:
                filters = rng.integers(-128, 128, size=(16, 18)).astype(np.int8)
                lib.oracle_lr_wiener_multi(*o_args, _p(filters), None, _p(sub), _p(noskip), base_q, edges, _p(ll),
                                           bd)
            else:
                filters = rng.integers(-500, 500, size=(64, 13)).astype(np.int16)
                lib.oracle_lr_wiener_multi(*o_args, None, _p(filters), _p(sub), _p(noskip), base_q, edges, _p(ll),
                                           bd)
            par.multi.filters = filters.ctypes.data
            par.multi.subclass_lut = sub.ctypes.data
            par.multi.noskip_mask = noskip.ctypes.data
            par.multi.base_q = base_q
        R._check(ctx.lib.rav2d_gpu_lr_wiener(ctx.h, kind, dev.at(2, 8), W * it_, dl.ptr, dev.at(0, 8),
                                             dev.at(h + 2, 8), w, h, C.byref(par), edges, _p(ll), bd, None),
                 "lr_wiener")
        got = dev.read()
        for d in [dev, dl] + keep:
            d.close()
        assert np.array_equal(got.astype(np.uint16), host), (it, w, h, edges)


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_gdf_prep_and_add(ctx, bd):
    rng = np.random.default_rng(300 + bd)
    lib = _olib()
    for it in range(10):
        w, h = [(64, 64), (64, 56), (32, 16), (16, 8), (64, 8)][it % 5]
        edges = int(rng.integers(0, 16)) if it else 15
        ref, qp = int(rng.integers(0, 6)), int(rng.integers(0, 6))
        img, left = _stripe(rng, w, h, bd)
        W = img.shape[1]
        dev, host = _gpu_and_host(ctx, img, bd)
        dl, hl = _gpu_and_host(ctx, left, bd)
        want = np.zeros((h, w), dtype=np.int8)
        lib.oracle_lr_gdf_prep(_p(want), w, _addr16(host, 2, 8), W, _p(hl), _addr16(host, 0, 8),
                               _addr16(host, h + 2, 8), w, h, ref, qp, edges, bd)
        err = ctx.device_array(np.zeros((h, w), dtype=np.int8))
        it_ = dev.dtype.itemsize
        R._check(ctx.lib.rav2d_gpu_gdf_prep(ctx.h, err.ptr, w, dev.at(2, 8), W * it_, dl.ptr, dev.at(0, 8),
                                            dev.at(h + 2, 8), w, h, ref, qp, edges, bd, None), "gdf_prep")
        got = err.read()
        assert np.array_equal(got, want), ("prep", it, ref, qp, edges)
        ll = _ll_rows(rng, h, w)
        scale = int(rng.integers(1, 5))
        lib.oracle_lr_gdf_add(_addr16(host, 2, 8), W, _p(want), w, w, h, scale, _p(ll), bd)
        R._check(ctx.lib.rav2d_gpu_gdf_add(ctx.h, dev.at(2, 8), W * it_, err.ptr, w, w, h, scale, _p(ll), bd, None),
                 "gdf_add")
        assert np.array_equal(dev.read().astype(np.uint16), host), ("add", it)
        for d in (dev, dl, err):
            d.close()


# chroma subsampling (sx, sy) of 4:2:0, 4:2:2 and 4:4:4
SUBSAMPLINGS = [(1, 1), (1, 0), (0, 0)]


def _grain_data(rng, lag=3):
    d = R.FilmGrainData()
    d.ar_coeff_lag = lag
    d.ar_coeff_shift = int(rng.integers(6, 10))
    d.grain_scale_shift = int(rng.integers(0, 4))
    d.scaling_shift = int(rng.integers(8, 12))
    for pl in range(3):
        for i in range(28):
            d.ar_coeffs[pl][i] = int(rng.integers(-128, 128)) // 2
    d.num_points[0] = 1
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("lag", [0, 1, 3])
def test_generate_grain(ctx, bd, lag):
    rng = np.random.default_rng(400 + bd + lag)
    L = O.load()
    for it in range(3):
        d = _grain_data(rng, lag)
        d.num_points[0] = it % 2
        seed = int(rng.integers(0, 1 << 16))
        want_y = np.zeros((73, 82), dtype=np.int16)
        L.oracle_generate_grain_y(O.ptr(want_y), C.byref(d), seed, bd)
        gy = ctx.device_array(np.zeros((73, 82), dtype=np.int16))
        R._check(ctx.lib.rav2d_gpu_generate_grain_y(ctx.h, gy.ptr, C.byref(d), seed, bd, None), "grain_y")
        assert np.array_equal(gy.read(), want_y), ("y", it)
        for (sx, sy) in SUBSAMPLINGS:
            th, tw = (38 if sy else 73), (44 if sx else 82)  # the chroma template's size
            for uv in range(2):
                want = np.zeros((73, 82), dtype=np.int16)
                L.oracle_generate_grain_uv(O.ptr(want), O.ptr(want_y), C.byref(d), seed, uv, sx, sy, bd)
                g = ctx.device_array(np.zeros((73, 82), dtype=np.int16))
                R._check(ctx.lib.rav2d_gpu_generate_grain_uv(ctx.h, g.ptr, gy.ptr, C.byref(d), seed, uv, sx, sy, bd,
                                                             None), "grain_uv")
                assert np.array_equal(g.read()[:th, :tw], want[:th, :tw]), ("uv", uv, it, sx, sy)
                g.close()
        gy.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("bs,ov", [(0, 1), (1, 1), (0, 0), (1, 0)])
@pytest.mark.parametrize("sx,sy", SUBSAMPLINGS, ids=["420", "422", "444"])
def test_fgy_fguv(ctx, bd, bs, ov, sx, sy):
    rng = np.random.default_rng(500 + bd * 8 + bs * 2 + ov + 1000 * (sx + sy))
    L = O.load()
    lib = _olib()
    for it in range(4):
        d = _grain_data(rng)
        d.block_size, d.overlap_flag = bs, ov
        d.clip_to_restricted_range = it & 1
        d.mc_identity = (it >> 1) & 1
        d.chroma_scaling_from_luma = 1 if it == 3 else 0
        for uv in range(2):
            d.uv_mult[uv], d.uv_luma_mult[uv] = int(rng.integers(-128, 128)), int(rng.integers(-128, 128))
            d.uv_offset[uv] = int(rng.integers(-256, 256))
        seed = int(rng.integers(0, 1 << 16))
        lut = np.zeros((73, 82), dtype=np.int16)
        L.oracle_generate_grain_y(O.ptr(lut), C.byref(d), seed, bd)
        scaling = rng.integers(0, 256, size=1 << bd).astype(np.uint8)
        bsz = 16 << bs
        pw = [200, 96, 248, 136][it]
        row_num = [0, 3, 1, 5][it]
        bh = [bsz, bsz, bsz - 8, 8][it]
        src = rng.integers(0, 1 << bd, size=(bh, pw))
        ds, hs = _gpu_and_host(ctx, src, bd)
        dd, hd = _gpu_and_host(ctx, np.zeros_like(src), bd)
        dsc, dlut = ctx.device_array(scaling), ctx.device_array(lut)
        lib.oracle_fgy_16(_p(hd), _p(hs), pw, C.byref(d), seed, pw, _p(scaling), _p(lut), bh, row_num, bd)
        it_ = ds.dtype.itemsize
        R._check(ctx.lib.rav2d_gpu_fgy_32x32xn(ctx.h, dd.ptr, ds.ptr, pw * it_, C.byref(d), seed, pw, dsc.ptr, dlut.ptr,
                                               bh, row_num, bd, None), "fgy")
        assert np.array_equal(dd.read().astype(np.uint16), hd), ("fgy", it)
        # chroma of the same band: bh >> sy rows, pw >> sx wide, luma = src
        cw, ch = pw >> sx, (bh + sy) >> sy
        csrc = rng.integers(0, 1 << bd, size=(ch, cw))
        dcs, hcs = _gpu_and_host(ctx, csrc, bd)
        dcd, hcd = _gpu_and_host(ctx, np.zeros_like(csrc), bd)
        for uv in range(2):
            L.oracle_generate_grain_uv(O.ptr(lut), O.ptr(np.zeros((73, 82), dtype=np.int16)), C.byref(d), seed, uv,
                                       sx, sy, bd)
            dlut.write(lut)
            lib.oracle_fguv_16(_p(hcd), _p(hcs), cw, C.byref(d), seed, cw, _p(scaling), _p(lut), ch, row_num, _p(hs),
                               pw, uv, d.mc_identity, sx, sy, bd)
            R._check(ctx.lib.rav2d_gpu_fguv_32x32xn(ctx.h, dcd.ptr, dcs.ptr, cw * it_, C.byref(d), seed, cw, dsc.ptr,
                                                    dlut.ptr, ch, row_num, ds.ptr, pw * it_, uv, d.mc_identity, sx, sy,
                                                    bd, None), "fguv")
            assert np.array_equal(dcd.read().astype(np.uint16), hcd), ("fguv", it, uv, sx, sy)
        for x in (ds, dd, dsc, dlut, dcs, dcd):
            x.close()
