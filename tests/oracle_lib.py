"""ctypes loader for the CPU oracle (oracle/liboracle.so): the checker the
parity tests compare the HIP path against."""
import ctypes as C
import os

import numpy as np

import rav2d_amd as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_lib = None


def load():
    global _lib
    if _lib is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so not built: run `python -m rav2d_amd.build`")
        lib = C.CDLL(path)
        P = C.c_void_p
        lib.oracle_filter_frame.restype = C.c_int
        lib.oracle_filter_frame.argtypes = [C.POINTER(R.FrameParams), P, C.c_size_t, P, C.c_size_t,
                                            P, P, P, C.c_ssize_t, C.c_ssize_t, C.c_uint32]
        lib.oracle_filter_frame_seg.restype = C.c_int
        lib.oracle_filter_frame_seg.argtypes = [C.POINTER(R.FrameParams), P, C.c_size_t, P, C.c_size_t,
                                                P, P, P, C.c_ssize_t, C.c_ssize_t, C.c_uint32,
                                                P, C.c_ssize_t, P, C.c_ssize_t]
        lib.oracle_apply_grain.restype = C.c_int
        lib.oracle_apply_grain.argtypes = [C.POINTER(R.FrameParams), P, P, P, P, P, P, C.c_ssize_t, C.c_ssize_t]
        lib.oracle_picture_has_grain.restype = C.c_int
        lib.oracle_picture_has_grain.argtypes = [C.POINTER(R.FrameParams)]
        lib.oracle_dq_lookup.restype = C.c_int
        lib.oracle_dq_lookup.argtypes = [C.c_int]
        lib.oracle_deblock_quant_thr.restype = C.c_uint32
        lib.oracle_deblock_quant_thr.argtypes = [C.c_int, C.c_int]
        lib.oracle_deblock_side_thr.restype = C.c_uint32
        lib.oracle_deblock_side_thr.argtypes = [C.c_int, C.c_int]
        P8 = C.POINTER(C.c_uint8)
        sig = {
            "oracle_filter_choice": (C.c_int, [P, C.c_int, C.c_ssize_t, C.c_ssize_t, C.c_ssize_t, C.c_int, C.c_int,
                                               C.c_uint32, C.c_uint32]),
            "oracle_edge_thr": (C.c_int, [C.c_int, C.c_int]),
            "oracle_cdef_constrain": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "oracle_cdef_adjust_strength": (C.c_int, [C.c_int, C.c_uint32]),
            "oracle_cdef_find_dir": (C.c_int, [P, C.c_int, C.c_ssize_t, C.c_int, C.POINTER(C.c_uint32)]),
            "oracle_ccso_score": (C.c_uint32, [C.c_int, C.c_int, C.c_uint32]),
            "oracle_ccso_prep_8": (None, [P, C.c_int, P, C.c_ssize_t, C.c_ssize_t, P, P, C.c_ssize_t, P,
                                          C.c_ssize_t, C.c_uint32, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, C.c_int]),
            "oracle_ccso_add_8": (None, [P, C.c_ssize_t, P, C.c_int, P, P, C.c_int, C.c_int, P]),
            "oracle_get_qval_given_tskip": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
            "oracle_backup_row_8": (None, [P, P, P, C.c_int, C.c_int, C.c_int]),
            "oracle_backup_row_lpf_8": (None, [P, P, C.c_int, C.c_int, C.c_int]),
            "oracle_backup_row_luma_8": (None, [P, P, C.c_ssize_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "oracle_compute_gradient_row_8": (None, [P, P, C.c_int, C.c_int, C.c_int, C.c_int]),
            "oracle_gdf_add_8": (None, [P, C.c_ssize_t, P, C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_int]),
            "oracle_grain_random": (C.c_uint32, [C.c_uint32, C.POINTER(C.c_uint32)]),
            "oracle_grain_round2": (C.c_int, [C.c_int, C.c_uint32]),
            "oracle_generate_scaling": (None, [C.c_int, P, C.c_int, P]),
            "oracle_generate_grain_y": (None, [P, C.POINTER(R.FilmGrainData), C.c_uint32, C.c_int]),
            "oracle_generate_grain_uv": (None, [P, P, C.POINTER(R.FilmGrainData), C.c_uint32, C.c_int, C.c_int,
                                                C.c_int, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def ptr(a):
    """ctypes address of a numpy array."""
    return a.ctypes.data_as(C.c_void_p)


def filter_frame(sf, inloop_mask=R.INLOOP_ALL, params=None, y_stride=None):
    """Run the oracle chain on copies of a SynthFrame's planes; returns [y, u, v].
    y_stride: samples per luma row of the oracle's picture (default: the
    padded plane's width); the reference's 4:4:4 chroma restoration reads its
    luma line store, whose rows have that length, past the frame width."""
    lib = load()
    # the planes inside their picture padding (8 rows below, 8 columns right),
    # which the deblocking reads where transforms overhang the frame edge
    padded = sf.padded_planes()
    if y_stride is not None and y_stride != padded[0].shape[1]:
        # rows of exactly y_stride samples (narrower than the padded plane only
        # for widths that are multiples of 128, which read no padding to the right)
        assert y_stride >= sf.width
        rows = np.zeros((padded[0].shape[0], y_stride), dtype=padded[0].dtype)
        n = min(y_stride, padded[0].shape[1])
        rows[:, :n] = padded[0][:, :n]
        padded[0] = rows
    bps = padded[0].itemsize
    p = params if params is not None else sf.params
    seg, seg_stride, seg_uv, uv_seg_stride = sf.segmaps()
    rc = lib.oracle_filter_frame_seg(C.byref(p), sf.mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr,
                                     padded[0].ctypes.data, padded[1].ctypes.data, padded[2].ctypes.data,
                                     padded[0].shape[1] * bps, padded[1].shape[1] * bps, inloop_mask,
                                     seg, seg_stride, seg_uv, uv_seg_stride)
    if rc != 0:
        raise RuntimeError("oracle_filter_frame failed: %d" % rc)
    return [a[:sf.plane_shape(pl)[0], :sf.plane_shape(pl)[1]].copy() for pl, a in enumerate(padded)]


def apply_grain(sf, planes, params=None):
    """Film grain of the oracle on in-loop output planes; returns the display planes."""
    import numpy as np
    lib = load()
    p = params if params is not None else sf.params
    out = [np.empty_like(a) for a in planes]
    bps = planes[0].itemsize
    rc = lib.oracle_apply_grain(C.byref(p), planes[0].ctypes.data, planes[1].ctypes.data, planes[2].ctypes.data,
                                out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data,
                                planes[0].shape[1] * bps, planes[1].shape[1] * bps)
    if rc != 0:
        raise RuntimeError("oracle_apply_grain failed: %d" % rc)
    return out
