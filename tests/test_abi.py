"""The C-ABI library loads, exports every entry point include/rav2d_gpu.h
declares, and its struct layouts match the Python mirrors (no compute)."""
import ctypes as C
import os
import re

import rav2d_amd as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
        names += re.findall(r"\b(rav2d_gpu_\w+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = R.load_gpu()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_mirrors():
    lib = R.load_gpu()
    assert lib.rav2d_gpu_struct_size(0) == R.AV2FILTER_SIZE
    assert lib.rav2d_gpu_struct_size(1) == R.AV2RESTORATION_SIZE
    assert lib.rav2d_gpu_struct_size(2) == C.sizeof(R.FrameParams)
    assert lib.rav2d_gpu_struct_size(3) == C.sizeof(R.FilmGrainData)
    assert lib.rav2d_gpu_struct_size(4) == C.sizeof(R.Planes)
    assert lib.rav2d_gpu_struct_size(5) == C.sizeof(R.BatchConfig)
    assert lib.rav2d_gpu_struct_size(6) == C.sizeof(R.ServeConfig)
    assert lib.rav2d_gpu_struct_size(7) == C.sizeof(R.ServeResult)
    assert lib.rav2d_gpu_struct_size(8) == C.sizeof(R.WienerParams)
    assert lib.rav2d_gpu_struct_size(9) == R.PACKED_FILTER_SIZE == 8004
    assert lib.rav2d_gpu_struct_size(99) == 0


def test_pack_filter_keeps_every_field_of_a_420_record():
    """rav2d_gpu_pack_filter is a host routine: each field of the device form
    equals the Av2Filter field it comes from (chroma cut to 32 x 32 units)."""
    import numpy as np
    sf = R.SynthFrame(600, 520, 10, 0, seed=3, mask_flags=R.MASK_TXPART | R.MASK_SUBPU, lossless_pct=10)
    try:
        raw = np.frombuffer(bytes(sf.mask_bytes()), np.uint8).reshape(sf.n_mask, R.AV2FILTER_SIZE)
        packed = R.pack_filter(sf.mask_ptr, sf.n_mask)
        assert packed.shape == (sf.n_mask, R.PACKED_FILTER_SIZE)

        def u16(a, off, shape):
            n = int(np.prod(shape))
            return a[off:off + 2 * n].copy().view(np.uint16).reshape(shape)

        for k in range(sf.n_mask):
            a, b = raw[k], packed[k]
            fa_y, fb_y = u16(a, 0, (2, 64, 5, 4)), u16(b, 0, (2, 64, 5, 4))
            np.testing.assert_array_equal(fa_y, fb_y)
            fa_uv, fb_uv = u16(a, 5120, (2, 64, 5, 4)), u16(b, 5120, (2, 32, 5, 2))
            np.testing.assert_array_equal(fa_uv[:, :32, :, :2], fb_uv)
            assert not fa_uv[:, 32:].any() and not fa_uv[:, :, :, 2:].any()  # nothing dropped
            # Av2Filter: qidx@10240 gdf@10272 cdef_idx@10288 ccso@10304 noskip@10308
            # lr_noskip@10564 ll_y@11076 ll_uv@11588
            np.testing.assert_array_equal(u16(a, 11076, (64, 4)), u16(b, 6400, (64, 4)))
            np.testing.assert_array_equal(u16(a, 11588, (64, 4))[:32], u16(b, 6912, (32, 4)))
            np.testing.assert_array_equal(u16(a, 10308, (32, 4)), u16(b, 7168, (32, 4)))
            np.testing.assert_array_equal(u16(a, 10564, (64, 4)), u16(b, 7424, (64, 4)))
            np.testing.assert_array_equal(u16(a, 10240, (16,)), u16(b, 7936, (16,)))
            assert bytes(a[10288:10304]) == bytes(b[7968:7984])    # cdef_idx
            assert bytes(a[10304:10307]) == bytes(b[7984:7987])    # ccso
            assert bytes(a[10272:10288]) == bytes(b[7988:8004])    # gdf
    finally:
        sf.close()


def test_version_string():
    assert b"gfx950" in R.load_gpu().rav2d_gpu_version()
