"""The filter tables in rav2d_amd/csrc/tables.h (extracted as text from the
reference's tables.rs / gdf_tables.rs): shapes, spot values the reference's
own tests and the AV1 specification fix, and the invariants the filters rely on."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load():
    txt = open(os.path.join(ROOT, "rav2d_amd", "csrc", "tables.h")).read()
    out = {}
    for m in re.finditer(r"static const (\w+) R2T_(\w+)((?:\[\d+\])+) = \{(.*?)\};", txt, re.S):
        dims = [int(d) for d in re.findall(r"\[(\d+)\]", m.group(3))]
        vals = [int(v) for v in re.findall(r"-?\d+", m.group(4))]
        out[m.group(2)] = np.array(vals, dtype=np.int64).reshape(dims)
    return out


T = _load()


def test_shapes():
    shapes = {
        "CDEF_DIRECTIONS": (12, 2), "CCSO_QUANT_SZ": (4, 4), "CCSO_OFFSET": (4, 8),
        "NS_WIENER_COEF_RANGE_Y": (16, 2), "NS_WIENER_COEF_RANGE_UV": (18, 2),
        "DEBLOCK_SIDE_THRESHOLDS": (296,), "PC_WIENER_FILTERS": (4, 64, 13),
        "PC_WIENER_SUB_CLASSIFY": (4, 256), "PC_WIENER_SUB_CLASSIFY_NS": (4, 7, 256),
        "GAUSSIAN_SEQUENCE": (2048,), "PC_WIENER_LUT_TO_CLASS": (4096,),
        "GDF_ALPHA": (3168,), "GDF_WEIGHT": (9504,), "GDF_BIAS": (36, 3),
        "GDF_INTRA_ERROR": (24576,), "GDF_INTER_ERROR": (30000,), "DQ_LOOKUP": (24,),
    }
    for name, shape in shapes.items():
        assert T[name].shape == shape, name
    # GDF: 6 ref kinds x 6 qp x 88 (22 features x 4 classes), x3 for the weights;
    # error tables 6 qp x 16^3 (intra) and 5 refs x 6 qp x 10^3 (inter)
    assert 6 * 6 * 88 == 3168 and 6 * 6 * 264 == 9504 and 6 * 16 ** 3 == 24576 and 5 * 6 * 10 ** 3 == 30000


def test_spot_values_from_reference_tests():
    assert list(T["CDEF_DIRECTIONS"][0]) == [12, 24]
    assert list(T["CCSO_OFFSET"][0]) == [0, 1, -1, 3, -3, 7, -7, -10]
    assert T["DEBLOCK_SIDE_THRESHOLDS"][0] == -16


def test_cdef_directions_are_the_av1_directions():
    # AV1 spec Cdef_Directions, (dy, dx) of the two taps of directions 0..7;
    # the table stores dy * 12 + dx, padded with directions 6, 7 before and 0, 1 after
    av1 = [[(-1, 1), (-2, 2)], [(0, 1), (-1, 2)], [(0, 1), (0, 2)], [(0, 1), (1, 2)],
           [(1, 1), (2, 2)], [(1, 0), (2, 1)], [(1, 0), (2, 0)], [(1, 0), (2, -1)]]
    for i in range(12):
        d = (i - 2) % 8
        assert [dy * 12 + dx for dy, dx in av1[d]] == list(T["CDEF_DIRECTIONS"][i]), i


def test_gaussian_sequence_is_the_av1_sequence():
    # first 16 entries of the AV1 spec's Gaussian_Sequence; all entries are
    # multiples of 4 and fit 12 bits signed
    g = T["GAUSSIAN_SEQUENCE"]
    assert list(g[:16]) == [56, 568, -180, 172, 124, -84, 172, -64, -900, 24, 820, 224, 1248, 996, 272, -8]
    assert np.all(g % 4 == 0)
    assert g.min() >= -2048 and g.max() < 2048
    assert abs(g.mean()) < 16  # zero-mean noise


def test_ccso_offsets_are_antisymmetric_scalings():
    off = T["CCSO_OFFSET"]
    for s in range(4):
        assert list(off[s]) == [v * (s + 1) for v in off[0]]
    assert all(off[0][2 * k - 1] == -off[0][2 * k] for k in range(1, 4))


def test_deblock_side_thresholds_non_decreasing():
    t = T["DEBLOCK_SIDE_THRESHOLDS"]
    assert np.all(np.diff(t) >= 0)


def test_pc_wiener_filters_preserve_dc():
    # out = (m * c[12] + sum_k (a_k + b_k) * c[k] + 64) >> 7: unit DC gain needs
    # 2 * sum(c[0..11]) + c[12] == 128 for every class filter
    f = T["PC_WIENER_FILTERS"]
    dc = 2 * f[:, :, :12].sum(axis=2) + f[:, :, 12]
    assert np.all(dc == 128)


def test_class_tables_index_in_range():
    assert T["PC_WIENER_LUT_TO_CLASS"].min() >= 0
    assert T["PC_WIENER_SUB_CLASSIFY"].max() < 64
    assert T["PC_WIENER_SUB_CLASSIFY_NS"].max() < 16
    lut = T["PC_WIENER_LUT_TO_CLASS"]
    assert lut.max() < T["PC_WIENER_SUB_CLASSIFY"].shape[1]


def test_ns_wiener_coefficient_ranges():
    for name, n in (("NS_WIENER_COEF_RANGE_Y", 16), ("NS_WIENER_COEF_RANGE_UV", 18)):
        r = T[name]  # [bits, min]: coefficient range [min, min + 2^bits - 1]
        assert r.shape == (n, 2)
        assert np.all((r[:, 0] >= 1) & (r[:, 0] <= 8)) and np.all(r[:, 1] < 0)
        assert np.all(r[:, 1] + (1 << r[:, 0]) - 1 > 0)  # ranges straddle zero
        assert np.array_equal(r[0::2], r[1::2])  # taps come in (point, mirror) pairs


def test_dq_lookup_is_one_octave():
    d = T["DQ_LOOKUP"]
    assert np.all(np.diff(d) > 0)
    assert d[0] * 2 > d[-1] and d[0] * 2 <= d[-1] + 4  # the next octave starts at 2 * d[0]
