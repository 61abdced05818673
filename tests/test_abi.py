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
    assert lib.rav2d_gpu_struct_size(99) == 0


def test_version_string():
    assert b"gfx950" in R.load_gpu().rav2d_gpu_version()
