"""rav2d_amd: MI355X (gfx950) post-reconstruction filter stage of the rav2d AV2 decoder.

Python bindings over the C ABI in ``include/rav2d_gpu.h`` (``librav2d_gpu.so``)
and over the seeded workload generator (``librav2d_synth.so``).  The compute
path is the HIP library; nothing here filters pixels on the CPU.  Loading the
GPU library fails loudly if it was not built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

RAV2D_OK = 0
RAV2D_EINVAL, RAV2D_ENOMEM, RAV2D_EDEVICE, RAV2D_EAGAIN, RAV2D_ENOENT, RAV2D_ECANCELED = -22, -12, -5, -11, -2, -125
FAIL_ALLOC, FAIL_LAUNCH = 0, 1


def inject_failure(what, count):
    """The next `count` library allocations (FAIL_ALLOC) or filter launches
    (FAIL_LAUNCH) fail on the host (rav2d_gpu_debug_inject_failure)."""
    _check(load_gpu().rav2d_gpu_debug_inject_failure(what, count), "inject_failure")
INLOOP_DEBLOCK, INLOOP_CDEF, INLOOP_CCSO, INLOOP_WIENER, INLOOP_GDF = 1, 2, 4, 8, 16
INLOOP_ALL = 31
PASS_DEBLOCK_V, PASS_DEBLOCK_H, PASS_CDEF, PASS_LR, PASS_GRAIN = range(5)
PIXEL_LAYOUT_I400, PIXEL_LAYOUT_I420, PIXEL_LAYOUT_I422, PIXEL_LAYOUT_I444 = range(4)


def chroma_shape(layout, width, height):
    """(rows, columns) of a chroma plane of a width x height frame: (0, 0) for
    I400, halved horizontally for 4:2:0 and 4:2:2, vertically for 4:2:0."""
    if layout == PIXEL_LAYOUT_I400:
        return (0, 0)
    ssx, ssy = int(layout != PIXEL_LAYOUT_I444), int(layout == PIXEL_LAYOUT_I420)
    return ((height + ssy) >> ssy, (width + ssx) >> ssx)

TOOL_DEBLOCK, TOOL_CDEF, TOOL_CCSO, TOOL_WIENER, TOOL_GDF, TOOL_GRAIN = 1, 2, 4, 8, 16, 32
TOOL_ALL = 63
# synthetic mask shapes: transform partitions (incl. H5/V5), sub-PU / TIP edges, segmentation
MASK_TXPART, MASK_SUBPU, MASK_SEG = 1, 2, 4


class FilmGrainData(C.Structure):
    _fields_ = [
        ("chroma_scaling_from_luma", C.c_int32),
        ("num_points", C.c_int32 * 3),
        ("points", (C.c_uint8 * 2 * 14) * 3),
        ("scaling_shift", C.c_int32),
        ("ar_coeff_lag", C.c_int32),
        ("ar_coeffs", (C.c_int8 * 28) * 3),
        ("ar_coeff_shift", C.c_uint64),
        ("grain_scale_shift", C.c_int32),
        ("uv_mult", C.c_int32 * 2),
        ("uv_luma_mult", C.c_int32 * 2),
        ("uv_offset", C.c_int32 * 2),
        ("overlap_flag", C.c_int32),
        ("clip_to_restricted_range", C.c_int32),
        ("mc_identity", C.c_int32),
        ("block_size", C.c_int32),
    ]


class CcsoPlane(C.Structure):
    _fields_ = [
        ("enabled", C.c_uint8), ("bo_only", C.c_uint8), ("scale_idx", C.c_uint8),
        ("quant_idx", C.c_uint8), ("ext_filter_support", C.c_uint8), ("edge_clf", C.c_uint8),
        ("max_band_log2", C.c_uint8), ("_pad", C.c_uint8), ("filter_off", C.c_uint8 * 64),
    ]


class RestorationPlane(C.Structure):
    _fields_ = [
        ("restoration_type", C.c_uint8), ("frame_filters_on", C.c_uint8),
        ("num_classes_idx", C.c_uint8), ("num_classes", C.c_uint8),
        ("filter", (C.c_int8 * 18) * 16),
    ]


class FrameParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("bitdepth", C.c_int32),
        ("layout", C.c_int32), ("sb128", C.c_int32), ("inloop_filters", C.c_uint32),
        ("seq_cdef", C.c_int32), ("cfl_ds_filter_index", C.c_int32),
        ("yac", C.c_int32), ("uac_delta", C.c_int32), ("vac_delta", C.c_int32),
        ("seg_enabled", C.c_int32), ("seg_delta_q", C.c_int16 * 8), ("seg_lossless", C.c_uint8 * 8),
        ("db_level_y", C.c_uint8 * 2), ("db_level_u", C.c_uint8), ("db_level_v", C.c_uint8),
        ("db_delta_q_y", C.c_int8 * 2), ("db_delta_q_u", C.c_int8), ("db_delta_q_v", C.c_int8),
        ("cdef_damping", C.c_uint8), ("cdef_n_strengths", C.c_uint8), ("cdef_on_skiptx", C.c_uint8),
        ("_pad0", C.c_uint8),
        ("cdef_y_strength", C.c_uint8 * 8), ("cdef_uv_strength", C.c_uint8 * 8),
        ("ccso", CcsoPlane * 3),
        ("restoration", RestorationPlane * 3),
        ("lr_unit_size", C.c_uint8 * 2), ("_pad1", C.c_uint8 * 2),
        ("gdf_enabled", C.c_int32), ("gdf_qp_idx", C.c_int32), ("gdf_scale", C.c_int32),
        ("gdf_ref_dst_idx", C.c_int32),
        ("grain_present", C.c_int32), ("grain_seed", C.c_uint32),
        ("grain", FilmGrainData),
    ]


class Planes(C.Structure):
    _fields_ = [("data", C.c_void_p * 3), ("stride", C.c_ssize_t * 2)]


NUM_PASSES = 5
PASS_NAMES = ("deblock_v", "deblock_h", "cdef_ccso", "loop_restoration", "film_grain")


class BatchConfig(C.Structure):
    _fields_ = [
        ("use_graph", C.c_int32),
        ("apply_grain", C.c_int32),
        ("grid", C.c_int32 * NUM_PASSES),
        ("xcd_remap", C.c_int32 * NUM_PASSES),
        ("deblock_rows", C.c_int32),
        ("mixed_depths", C.c_int32),
    ]

    def as_dict(self):
        return {"use_graph": self.use_graph, "apply_grain": self.apply_grain,
                "grid": list(self.grid), "xcd_remap": list(self.xcd_remap),
                "deblock_rows": self.deblock_rows, "mixed_depths": self.mixed_depths}


class ServeConfig(C.Structure):
    _fields_ = [
        ("max_batch", C.c_int32),
        ("max_inflight", C.c_int32),
        ("apply_grain", C.c_int32),
        ("checksum", C.c_int32),
    ]


class ServeResult(C.Structure):
    _fields_ = [
        ("tag", C.c_uint64),
        ("t_submit", C.c_double),
        ("t_launch", C.c_double),
        ("t_done", C.c_double),
        ("checksum", C.c_uint64 * 2),
        ("status", C.c_int32),
        ("batch_frames", C.c_int32),
    ]


class SynthConfig(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("bitdepth", C.c_int32), ("sb128", C.c_int32),
        ("seed", C.c_uint64), ("tools", C.c_uint32), ("lossless_pct", C.c_int32),
        ("skip_pct", C.c_int32), ("pc_gdf_pct", C.c_int32), ("mask_flags", C.c_uint32),
        ("monochrome", C.c_int32), ("chroma_format", C.c_int32),
    ]


AV2FILTER_SIZE = 12100
AV2RESTORATION_SIZE = 24624
PACKED_FILTER_SIZE = 8004  # Rav2dPackedFilter, the device form of an Av2Filter record
PICTURE_PAD = 8  # RAV2D_PICTURE_PAD: rows / columns of picture padding the deblocking reads

_gpu = None
_synth = None


def _declare(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def gpu_lib_path():
    return os.path.join(HERE, "librav2d_gpu.so")


def load_gpu():
    """Load the HIP library; raises if it was not built (no CPU fallback exists)."""
    global _gpu
    if _gpu is not None:
        return _gpu
    path = gpu_lib_path()
    if not os.path.exists(path):
        raise RuntimeError("librav2d_gpu.so not built: run `python -m rav2d_amd.build`")
    lib = C.CDLL(path)
    P = C.c_void_p
    _declare(lib, "rav2d_gpu_version", C.c_char_p, [])
    _declare(lib, "rav2d_gpu_last_error", C.c_char_p, [])
    _declare(lib, "rav2d_gpu_ctx_create", C.c_int, [C.c_int, C.c_uint64, C.c_uint32, C.POINTER(P)])
    _declare(lib, "rav2d_gpu_ctx_destroy", C.c_int, [P])
    _declare(lib, "rav2d_gpu_ctx_synchronize", C.c_int, [P])
    _declare(lib, "rav2d_gpu_ctx_bytes_in_use", C.c_uint64, [P])
    _declare(lib, "rav2d_gpu_frame_create", C.c_int, [P, C.POINTER(FrameParams), C.POINTER(P)])
    _declare(lib, "rav2d_gpu_frame_destroy", C.c_int, [P])
    _declare(lib, "rav2d_gpu_frame_set_side_info", C.c_int, [P, P, C.c_size_t, P, C.c_size_t, P])
    _declare(lib, "rav2d_gpu_frame_set_segmap", C.c_int, [P, P, C.c_ssize_t, P, C.c_ssize_t, P])
    _declare(lib, "rav2d_gpu_frame_set_side_info_packed", C.c_int, [P, P, C.c_size_t, P, C.c_size_t, P])
    _declare(lib, "rav2d_gpu_pack_filter", C.c_int, [P, C.c_size_t, P])
    _declare(lib, "rav2d_gpu_frame_set_params", C.c_int, [P, C.POINTER(FrameParams)])
    _declare(lib, "rav2d_gpu_frame_planes", C.c_int, [P, C.c_int, C.POINTER(Planes)])
    _declare(lib, "rav2d_gpu_frame_upload", C.c_int, [P, P, P, P, P])
    _declare(lib, "rav2d_gpu_frame_upload_padded", C.c_int, [P, C.POINTER(Planes), C.c_int, C.c_int, P])
    _declare(lib, "rav2d_gpu_frame_download", C.c_int, [P, C.c_int, P, P, P, P])
    _declare(lib, "rav2d_gpu_filter_frame", C.c_int, [P, P, P])
    _declare(lib, "rav2d_gpu_apply_grain", C.c_int, [P, P, P])
    _declare(lib, "rav2d_gpu_run_pass", C.c_int, [P, P, C.c_int, P])
    _declare(lib, "rav2d_gpu_timer_start", C.c_int, [P])
    _declare(lib, "rav2d_gpu_timer_stop", C.c_int, [P, C.POINTER(C.c_float)])
    _declare(lib, "rav2d_gpu_struct_size", C.c_size_t, [C.c_int])
    _declare(lib, "rav2d_gpu_device_count", C.c_int, [])
    _declare(lib, "rav2d_gpu_set_profiling", C.c_int, [P, C.c_int])
    _declare(lib, "rav2d_gpu_pass_times", C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int])
    _declare(lib, "rav2d_gpu_batch_default_config", None, [C.POINTER(BatchConfig)])
    _declare(lib, "rav2d_gpu_batch_create", C.c_int,
             [P, C.POINTER(P), C.c_int, C.POINTER(BatchConfig), C.POINTER(P)])
    _declare(lib, "rav2d_gpu_batch_destroy", C.c_int, [P])
    _declare(lib, "rav2d_gpu_batch_run", C.c_int, [P, P])
    _declare(lib, "rav2d_gpu_batch_get_config", C.c_int, [P, C.POINTER(BatchConfig)])
    _declare(lib, "rav2d_gpu_batch_set_config", C.c_int, [P, C.POINTER(BatchConfig)])
    _declare(lib, "rav2d_gpu_batch_autotune", C.c_int, [P, C.c_int, C.POINTER(BatchConfig)])
    _declare(lib, "rav2d_gpu_batch_info", C.c_int,
             [P, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int)])
    _declare(lib, "rav2d_gpu_batch_pixels", C.c_int64, [P])
    _declare(lib, "rav2d_gpu_server_default_config", None, [C.POINTER(ServeConfig)])
    _declare(lib, "rav2d_gpu_server_create", C.c_int, [P, C.POINTER(ServeConfig), C.POINTER(P)])
    _declare(lib, "rav2d_gpu_server_destroy", C.c_int, [P])
    _declare(lib, "rav2d_gpu_server_clock", C.c_double, [P])
    _declare(lib, "rav2d_gpu_server_submit", C.c_int, [P, P, C.c_uint64, C.c_double])
    _declare(lib, "rav2d_gpu_server_poll", C.c_int, [P, C.POINTER(ServeResult), C.c_int, C.c_int])
    _declare(lib, "rav2d_gpu_server_drain", C.c_int, [P])
    _declare(lib, "rav2d_gpu_server_pause", C.c_int, [P, C.c_int])
    _declare(lib, "rav2d_gpu_server_cancel", C.c_int, [P, C.c_uint64])
    _declare(lib, "rav2d_gpu_debug_inject_failure", C.c_int, [C.c_int, C.c_int])
    _declare(lib, "rav2d_gpu_server_simulate", C.c_int,
             [P, C.POINTER(P), C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint64, C.POINTER(ServeResult),
              C.POINTER(C.c_int32), C.POINTER(C.c_int32)])
    _declare(lib, "rav2d_gpu_frame_checksum", C.c_int, [P, C.c_int, C.POINTER(C.c_uint64)])
    _declare(lib, "rav2d_gpu_bandwidth_probe", C.c_int, [P, C.c_uint64, C.c_int, C.POINTER(C.c_double)])
    _declare(lib, "rav2d_gpu_picture_alloc", C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Planes)])
    _declare(lib, "rav2d_gpu_picture_free", C.c_int, [P, C.POINTER(Planes)])
    _declare(lib, "rav2d_gpu_filter_planes", C.c_int, [P, P, C.POINTER(Planes), C.POINTER(Planes), P])
    _declare(lib, "rav2d_gpu_apply_grain_planes", C.c_int,
             [P, C.POINTER(Planes), C.POINTER(Planes), C.c_int, C.c_int, C.c_int, C.c_int,
              C.POINTER(FilmGrainData), C.c_uint32, P])
    _declare(lib, "rav2d_gpu_apply_grain_planes_noalloc", C.c_int,
             [P, C.POINTER(Planes), C.POINTER(Planes), C.c_int, C.c_int, C.c_int, C.c_int,
              C.POINTER(FilmGrainData), C.c_uint32, P])
    _declare(lib, "rav2d_gpu_alloc", C.c_int, [P, C.c_size_t, C.POINTER(P)])
    _declare(lib, "rav2d_gpu_free", C.c_int, [P, P])
    _declare(lib, "rav2d_gpu_copy", C.c_int, [P, P, P, C.c_size_t, P])
    # per-family DSP entries
    I, U32, SS = C.c_int, C.c_uint32, C.c_ssize_t
    _declare(lib, "rav2d_gpu_deblock_sb", I, [P, I, I, P, SS, P, P, P, P, I, I, P])
    _declare(lib, "rav2d_gpu_cdef_dir", I, [P, P, SS, I, P, P, P])
    _declare(lib, "rav2d_gpu_cdef_filter", I, [P, P, SS, P, P, P, I, I, I, I, I, I, I, I, P])
    _declare(lib, "rav2d_gpu_ccso_prep", I, [P, P, SS, P, SS, P, P, P, I, I, I, U32, I, I, I, I, I, I, I, P])
    _declare(lib, "rav2d_gpu_ccso_add", I, [P, P, SS, P, SS, P, P, I, I, P, I, P])
    _declare(lib, "rav2d_gpu_lr_wiener", I, [P, I, P, SS, P, P, P, I, I, P, C.c_uint8, P, I, P])
    _declare(lib, "rav2d_gpu_gdf_prep", I, [P, P, SS, P, SS, P, P, P, I, I, I, I, C.c_uint8, I, P])
    _declare(lib, "rav2d_gpu_gdf_add", I, [P, P, SS, P, SS, I, I, I, P, I, P])
    _declare(lib, "rav2d_gpu_generate_grain_y", I, [P, P, C.POINTER(FilmGrainData), U32, I, P])
    _declare(lib, "rav2d_gpu_generate_grain_uv", I, [P, P, P, C.POINTER(FilmGrainData), U32, I, I, I, I, P])
    _declare(lib, "rav2d_gpu_fgy_32x32xn", I, [P, P, P, SS, C.POINTER(FilmGrainData), U32, I, P, P, I, I, I, P])
    _declare(lib, "rav2d_gpu_fguv_32x32xn", I, [P, P, P, SS, C.POINTER(FilmGrainData), U32, I, P, P, I, I, P, SS,
                                                I, I, I, I, I, P])
    _gpu = lib
    return lib


def load_synth():
    global _synth
    if _synth is not None:
        return _synth
    path = os.path.join(HERE, "librav2d_synth.so")
    if not os.path.exists(path):
        raise RuntimeError("librav2d_synth.so not built: run `python -m rav2d_amd.build`")
    lib = C.CDLL(path)
    P = C.c_void_p
    _declare(lib, "r2s_frame_create", P, [C.POINTER(SynthConfig)])
    _declare(lib, "r2s_frame_destroy", None, [P])
    _declare(lib, "r2s_params", C.POINTER(FrameParams), [P])
    _declare(lib, "r2s_plane", P, [P, C.c_int])
    _declare(lib, "r2s_plane_width", C.c_int32, [P, C.c_int])
    _declare(lib, "r2s_plane_height", C.c_int32, [P, C.c_int])
    _declare(lib, "r2s_plane_bytes", C.c_size_t, [P, C.c_int])
    _declare(lib, "r2s_mask", P, [P, C.POINTER(C.c_size_t)])
    _declare(lib, "r2s_lr", P, [P, C.POINTER(C.c_size_t)])
    _declare(lib, "r2s_segmap", P, [P, C.c_int, C.POINTER(C.c_ssize_t)])
    _declare(lib, "r2s_plane_padded", P, [P, C.c_int, C.POINTER(C.c_ssize_t)])
    _synth = lib
    return lib


def device_count():
    """HIP devices visible to the library (0 without a GPU)."""
    return load_gpu().rav2d_gpu_device_count()


class Rav2dError(RuntimeError):
    pass


def _check(rc, what):
    if rc != RAV2D_OK:
        msg = load_gpu().rav2d_gpu_last_error()
        raise Rav2dError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


def pack_filter(mask_ptr, n):
    """Av2Filter records at mask_ptr -> their device form (rav2d_gpu_pack_filter),
    as an (n, PACKED_FILTER_SIZE) uint8 array."""
    out = np.zeros((n, PACKED_FILTER_SIZE), np.uint8)
    _check(load_gpu().rav2d_gpu_pack_filter(mask_ptr, n, out.ctypes.data), "pack_filter")
    return out


class SynthFrame:
    """A seeded synthetic reconstructed frame plus its filter side info."""

    def __init__(self, width, height, bitdepth=10, sb128=0, seed=1, tools=TOOL_ALL,
                 lossless_pct=0, skip_pct=30, pc_gdf_pct=50, mask_flags=0, monochrome=False, layout=None):
        """layout: PIXEL_LAYOUT_I420 (default), _I422, _I444 or _I400 (the same as
        monochrome=True).  The luma of every layout equals that of the 4:2:0
        frame of the same seed."""
        if layout == PIXEL_LAYOUT_I400:
            monochrome = True
        fmt = {None: 0, PIXEL_LAYOUT_I400: 0, PIXEL_LAYOUT_I420: 0, PIXEL_LAYOUT_I422: 1, PIXEL_LAYOUT_I444: 2}[layout]
        lib = load_synth()
        cfg = SynthConfig(width, height, bitdepth, sb128, seed, tools, lossless_pct, skip_pct, pc_gdf_pct,
                          mask_flags, 1 if monochrome else 0, fmt)
        self._lib = lib
        self._h = lib.r2s_frame_create(C.byref(cfg))
        if not self._h:
            raise ValueError("invalid synthetic frame config")
        self.width, self.height, self.bitdepth = width, height, bitdepth
        self.monochrome = bool(monochrome)
        self.params = lib.r2s_params(self._h).contents
        self.layout = self.params.layout
        n = C.c_size_t()
        self.mask_ptr = lib.r2s_mask(self._h, C.byref(n))
        self.n_mask = n.value
        self.lr_ptr = lib.r2s_lr(self._h, C.byref(n))
        self.n_lr = n.value

    def plane_ptr(self, pl):
        return self._lib.r2s_plane(self._h, pl)

    def plane_shape(self, pl):
        return (self._lib.r2s_plane_height(self._h, pl), self._lib.r2s_plane_width(self._h, pl))

    @property
    def dtype(self):
        return np.uint16 if self.bitdepth > 8 else np.uint8

    def planes(self):
        """Copies of the three input planes as numpy arrays."""
        out = []
        for pl in range(3):
            h, w = self.plane_shape(pl)
            n = h * w * np.dtype(self.dtype).itemsize
            if n == 0:  # no chroma planes (monochrome)
                out.append(np.zeros((h, w), dtype=self.dtype))
                continue
            buf = (C.c_uint8 * n).from_address(self.plane_ptr(pl))
            out.append(np.frombuffer(buf, dtype=self.dtype).reshape(h, w).copy())
        return out

    def segmaps(self):
        """(luma map, stride, chroma map, stride) of a segmented frame, else Nones."""
        st, uvst = C.c_ssize_t(), C.c_ssize_t()
        y = self._lib.r2s_segmap(self._h, 0, C.byref(st))
        uv = self._lib.r2s_segmap(self._h, 1, C.byref(uvst))
        return y, st.value, uv, uvst.value

    def padded_planes(self):
        """Copies of the planes with 8 rows below / 8 columns right of picture
        padding (a decoder's reconstruction of the blocks overhanging the frame)."""
        out = []
        for pl in range(3):
            h, w = self.plane_shape(pl)
            if h * w == 0:  # no chroma planes (monochrome)
                out.append(np.zeros((0, 0), dtype=self.dtype))
                continue
            st = C.c_ssize_t()
            ptr = self._lib.r2s_plane_padded(self._h, pl, C.byref(st))
            isz = np.dtype(self.dtype).itemsize
            buf = (C.c_uint8 * (st.value * (h + PICTURE_PAD))).from_address(ptr)
            out.append(np.frombuffer(buf, dtype=self.dtype).reshape(h + PICTURE_PAD, st.value // isz).copy())
        return out

    def mask_bytes(self):
        return (C.c_uint8 * (self.n_mask * AV2FILTER_SIZE)).from_address(self.mask_ptr)

    def close(self):
        if self._h:
            self._lib.r2s_frame_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuContext:
    def __init__(self, device=0, hbm_budget=0):
        self.lib = load_gpu()
        h = C.c_void_p()
        _check(self.lib.rav2d_gpu_ctx_create(device, hbm_budget, 0, C.byref(h)), "ctx_create")
        self.h = h

    def synchronize(self):
        _check(self.lib.rav2d_gpu_ctx_synchronize(self.h), "synchronize")

    def bytes_in_use(self):
        return self.lib.rav2d_gpu_ctx_bytes_in_use(self.h)

    def set_profiling(self, on):
        _check(self.lib.rav2d_gpu_set_profiling(self.h, 1 if on else 0), "set_profiling")

    def pass_times(self, reset=True):
        """{pass index: (total ms, launches)} since the last reset."""
        ms = (C.c_double * 5)()
        cnt = (C.c_int64 * 5)()
        _check(self.lib.rav2d_gpu_pass_times(self.h, ms, cnt, 1 if reset else 0), "pass_times")
        return {i: (ms[i], cnt[i]) for i in range(5)}

    def timer_start(self):
        _check(self.lib.rav2d_gpu_timer_start(self.h), "timer_start")

    def timer_stop(self):
        ms = C.c_float()
        _check(self.lib.rav2d_gpu_timer_stop(self.h, C.byref(ms)), "timer_stop")
        return ms.value

    def device_array(self, host):
        """A DeviceArray holding a copy of the numpy array `host`."""
        return DeviceArray(self, host)

    def bandwidth_probe(self, nbytes=256 << 20, iters=20):
        """GB/s of a plain streaming copy (read + write bytes)."""
        g = C.c_double()
        _check(self.lib.rav2d_gpu_bandwidth_probe(self.h, nbytes, iters, C.byref(g)), "bandwidth_probe")
        return g.value

    def close(self):
        if self.h:
            self.lib.rav2d_gpu_ctx_destroy(self.h)
            self.h = None


class DeviceArray:
    """Device memory (rav2d_gpu_alloc) shaped like a numpy array: `ptr` is the
    device address of element 0; `read()` copies it back."""

    def __init__(self, ctx, host):
        self.ctx = ctx
        self.lib = ctx.lib
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = max(int(host.nbytes), 1)
        p = C.c_void_p()
        _check(self.lib.rav2d_gpu_alloc(ctx.h, self.nbytes, C.byref(p)), "alloc")
        self.ptr = p.value
        self.write(host)

    def write(self, host):
        h = np.ascontiguousarray(host, dtype=self.dtype)
        _check(self.lib.rav2d_gpu_copy(self.ctx.h, self.ptr, h.ctypes.data, h.nbytes, None), "copy")

    def read(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.lib.rav2d_gpu_copy(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes, None), "copy")
        return out

    def at(self, *index):
        """Device address of element `index`."""
        off = int(np.ravel_multi_index(index, self.shape)) * self.dtype.itemsize if index else 0
        return self.ptr + off

    def close(self):
        if self.ptr:
            self.lib.rav2d_gpu_free(self.ctx.h, self.ptr)
            self.ptr = None


class Picture:
    """A device picture from rav2d_gpu_picture_alloc (the library's plane
    layout: the strides rav2d_gpu_frame_planes reports for that size)."""

    def __init__(self, ctx, width, height, bitdepth, layout=1):
        self.ctx, self.lib = ctx, ctx.lib
        self.width, self.height, self.bitdepth = width, height, bitdepth
        self.layout = layout
        self.planes = Planes()
        _check(self.lib.rav2d_gpu_picture_alloc(ctx.h, width, height, bitdepth, layout, C.byref(self.planes)),
               "picture_alloc")

    @property
    def dtype(self):
        return np.uint16 if self.bitdepth > 8 else np.uint8

    def _shape(self, pl):
        if pl == 0:
            return (self.height, self.width)
        return chroma_shape(self.layout, self.width, self.height)

    def write(self, planes):
        """Upload three numpy planes (rows of any stride are packed here); planes
        with picture padding (up to PICTURE_PAD more rows and columns) fill the
        picture's padding too.  Monochrome pictures take the luma plane only."""
        for pl, a in enumerate(planes):
            if not self.planes.data[pl]:
                continue
            h, w = self._shape(pl)
            s = self.planes.stride[1 if pl else 0]
            cols = s // np.dtype(self.dtype).itemsize
            buf = np.zeros((h + PICTURE_PAD, cols), dtype=self.dtype)
            ah, aw = min(a.shape[0], h + PICTURE_PAD), min(a.shape[1], cols)
            buf[:ah, :aw] = a[:ah, :aw]
            _check(self.lib.rav2d_gpu_copy(self.ctx.h, self.planes.data[pl], buf.ctypes.data, buf.nbytes, None),
                   "copy")

    def read(self):
        out = []
        for pl in range(3):
            h, w = self._shape(pl)
            if not self.planes.data[pl]:
                out.append(np.zeros((h, w), dtype=self.dtype))
                continue
            s = self.planes.stride[1 if pl else 0]
            buf = np.empty((h, s // np.dtype(self.dtype).itemsize), dtype=self.dtype)
            _check(self.lib.rav2d_gpu_copy(self.ctx.h, buf.ctypes.data, self.planes.data[pl], buf.nbytes, None),
                   "copy")
            out.append(buf[:, :w].copy())
        return out

    def close(self):
        if self.planes.data[0]:
            self.lib.rav2d_gpu_picture_free(self.ctx.h, C.byref(self.planes))


class WienerParamsSingle(C.Structure):
    _fields_ = [("filter", C.c_void_p), ("luma", C.c_void_p), ("luma_top", C.c_void_p),
                ("luma_bottom", C.c_void_p), ("stride", C.c_ssize_t), ("ss_ver", C.c_int32),
                ("ss_hor", C.c_int32), ("ds_flt", C.c_int32)]


class WienerParamsMulti(C.Structure):
    _fields_ = [("filters", C.c_void_p), ("subclass_lut", C.c_void_p), ("noskip_mask", C.c_void_p),
                ("base_q", C.c_int32)]


class WienerParams(C.Union):
    _fields_ = [("single", WienerParamsSingle), ("multi", WienerParamsMulti)]


class GpuFrame:
    """A device-resident frame: input planes, output planes, display planes, side info."""

    def __init__(self, ctx, params):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        _check(self.lib.rav2d_gpu_frame_create(ctx.h, C.byref(params), C.byref(h)), "frame_create")
        self.h = h
        self.width, self.height, self.bitdepth = params.width, params.height, params.bitdepth
        self.layout = params.layout
        self.monochrome = params.layout == PIXEL_LAYOUT_I400

    @classmethod
    def from_synth(cls, ctx, sf):
        f = cls(ctx, sf.params)
        f.set_side_info(sf.mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr)
        if sf.params.seg_enabled:
            f.set_segmap(*sf.segmaps())
        f.upload_padded(sf)
        return f

    def upload_padded(self, sf):
        """The SynthFrame's planes with their picture padding (8 rows and columns)."""
        src = Planes()
        for pl in range(1 if self.monochrome else 3):
            st = C.c_ssize_t()
            src.data[pl] = sf._lib.r2s_plane_padded(sf._h, pl, C.byref(st))
            src.stride[1 if pl else 0] = st.value
        _check(self.lib.rav2d_gpu_frame_upload_padded(self.h, C.byref(src), PICTURE_PAD, PICTURE_PAD, None),
               "upload_padded")

    def set_side_info(self, mask_ptr, n_mask, lr_ptr, n_lr):
        _check(self.lib.rav2d_gpu_frame_set_side_info(self.h, mask_ptr, n_mask, lr_ptr, n_lr, None),
               "set_side_info")

    def set_params(self, params):
        _check(self.lib.rav2d_gpu_frame_set_params(self.h, C.byref(params)), "set_params")

    def set_segmap(self, segmap, stride, segmap_uv, uv_stride):
        """Segment maps (pointers or None, strides in bytes) of a segmented frame."""
        _check(self.lib.rav2d_gpu_frame_set_segmap(self.h, segmap, stride, segmap_uv, uv_stride, None),
               "set_segmap")

    def filter(self):
        _check(self.lib.rav2d_gpu_filter_frame(self.ctx.h, self.h, None), "filter_frame")

    def apply_grain(self):
        _check(self.lib.rav2d_gpu_apply_grain(self.ctx.h, self.h, None), "apply_grain")

    def run_pass(self, upto):
        _check(self.lib.rav2d_gpu_run_pass(self.ctx.h, self.h, upto, None), "run_pass")

    def download(self, which=1):
        dt = np.uint16 if self.bitdepth > 8 else np.uint8
        c = chroma_shape(self.layout, self.width, self.height)
        arrs = [np.empty(s, dtype=dt) for s in [(self.height, self.width), c, c]]
        ptrs = [a.ctypes.data if a.size else None for a in arrs]
        _check(self.lib.rav2d_gpu_frame_download(self.h, which, ptrs[0], ptrs[1], ptrs[2], None), "download")
        return arrs

    def checksum(self, which=1):
        out = C.c_uint64()
        _check(self.lib.rav2d_gpu_frame_checksum(self.h, which, C.byref(out)), "frame_checksum")
        return out.value

    def close(self):
        if self.h:
            self.lib.rav2d_gpu_frame_destroy(self.h)
            self.h = None


def default_batch_config():
    cfg = BatchConfig()
    load_gpu().rav2d_gpu_batch_default_config(C.byref(cfg))
    return cfg


class GpuBatch:
    """Many frames filtered with one launch per pass (rav2d_gpu_batch_*)."""

    def __init__(self, ctx, frames, config=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.frames = list(frames)
        arr = (C.c_void_p * len(self.frames))(*[f.h.value for f in self.frames])
        h = C.c_void_p()
        cfg = config if config is not None else default_batch_config()
        _check(self.lib.rav2d_gpu_batch_create(ctx.h, arr, len(self.frames), C.byref(cfg), C.byref(h)),
               "batch_create")
        self.h = h

    def run(self):
        _check(self.lib.rav2d_gpu_batch_run(self.h, None), "batch_run")

    def config(self):
        cfg = BatchConfig()
        _check(self.lib.rav2d_gpu_batch_get_config(self.h, C.byref(cfg)), "batch_get_config")
        return cfg

    def set_config(self, cfg):
        _check(self.lib.rav2d_gpu_batch_set_config(self.h, C.byref(cfg)), "batch_set_config")

    def autotune(self, iters=3):
        cfg = BatchConfig()
        _check(self.lib.rav2d_gpu_batch_autotune(self.h, iters, C.byref(cfg)), "batch_autotune")
        return cfg

    def info(self, p):
        items, nbytes, nf = C.c_int64(), C.c_int64(), C.c_int()
        _check(self.lib.rav2d_gpu_batch_info(self.h, p, C.byref(items), C.byref(nbytes), C.byref(nf)),
               "batch_info")
        return items.value, nbytes.value, nf.value

    def pixels(self):
        return self.lib.rav2d_gpu_batch_pixels(self.h)

    def close(self):
        if self.h:
            self.lib.rav2d_gpu_batch_destroy(self.h)
            self.h = None


def plane_checksum(planes):
    """Host restatement of rav2d_gpu_frame_checksum over numpy planes [y, u, v]."""
    M = np.uint64((1 << 64) - 1)
    total = 0
    for pl, a in enumerate(planes):
        h, w = a.shape
        y = np.arange(h, dtype=np.uint64)[:, None]
        x = np.arange(w, dtype=np.uint64)[None, :]
        z = (np.uint64(pl) << np.uint64(60)) | (y << np.uint64(40)) | (x << np.uint64(20)) | a.astype(np.uint64)
        with np.errstate(over="ignore"):
            z = z + np.uint64(0x9e3779b97f4a7c15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
            z = z ^ (z >> np.uint64(31))
        total = (total + int(z.sum(dtype=np.uint64))) & int(M)
    return total


def default_serve_config():
    cfg = ServeConfig()
    load_gpu().rav2d_gpu_server_default_config(C.byref(cfg))
    return cfg


class GpuServer:
    """Continuous-batching request loop (rav2d_gpu_server_*)."""

    def __init__(self, ctx, config=None):
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = config if config is not None else default_serve_config()
        h = C.c_void_p()
        _check(self.lib.rav2d_gpu_server_create(ctx.h, C.byref(cfg), C.byref(h)), "server_create")
        self.h = h

    def clock(self):
        return self.lib.rav2d_gpu_server_clock(self.h)

    def submit(self, frame, tag, t_arrival=-1.0):
        _check(self.lib.rav2d_gpu_server_submit(self.h, frame.h, tag, t_arrival), "server_submit")

    def poll(self, max_results=256, wait_ms=0):
        buf = (ServeResult * max_results)()
        n = self.lib.rav2d_gpu_server_poll(self.h, buf, max_results, wait_ms)
        if n < 0:
            _check(n, "server_poll")
        return [buf[i] for i in range(n)]

    def drain(self):
        _check(self.lib.rav2d_gpu_server_drain(self.h), "server_drain")

    def pause(self, on=True):
        _check(self.lib.rav2d_gpu_server_pause(self.h, 1 if on else 0), "server_pause")

    def cancel(self, tag):
        """True if the queued job was cancelled, False if no queued job has the tag."""
        rc = self.lib.rav2d_gpu_server_cancel(self.h, tag)
        if rc == RAV2D_ENOENT:
            return False
        _check(rc, "server_cancel")
        return True

    def simulate(self, stream_frames, slots, fps, n_jobs, seed):
        """stream_frames: list (per stream) of `slots` GpuFrames.  Returns
        (results, job_stream, job_slot) as numpy arrays / ServeResult list."""
        n_streams = len(stream_frames)
        flat = [f for fs in stream_frames for f in fs]
        assert all(len(fs) == slots for fs in stream_frames)
        arr = (C.c_void_p * len(flat))(*[f.h.value for f in flat])
        res = (ServeResult * n_jobs)()
        js = np.zeros(n_jobs, dtype=np.int32)
        jl = np.zeros(n_jobs, dtype=np.int32)
        _check(self.lib.rav2d_gpu_server_simulate(
            self.h, arr, n_streams, slots, fps, n_jobs, seed, res,
            js.ctypes.data_as(C.POINTER(C.c_int32)), jl.ctypes.data_as(C.POINTER(C.c_int32))), "server_simulate")
        return res, js, jl

    def close(self):
        if self.h:
            self.lib.rav2d_gpu_server_destroy(self.h)
            self.h = None
