#!/usr/bin/env python3
"""Generate ``tables.h`` (constant data for the post-reconstruction filters).

The filter stage must reproduce rav2d bit for bit, so its lookup tables must
hold exactly the reference's values.  This script is the only place that reads
the reference checkout: it parses the Rust array literals by name, flattens the
integers, checks each table's length against its declared shape, and writes
them as plain C arrays.  The output is committed; the build never runs this.

Usage:  python3 extract_tables.py /root/reference  > tables.h

Tables (source file:line in the reference checkout):
  crates/rav2d/src/tables.rs:385   CDEF_DIRECTIONS[12][2]
  crates/rav2d/src/tables.rs:435   CCSO_QUANT_SZ[4][4]
  crates/rav2d/src/tables.rs:442   CCSO_OFFSET[4][8]
  crates/rav2d/src/tables.rs:471   NS_WIENER_COEF_RANGE_Y[16][2]
  crates/rav2d/src/tables.rs:490   NS_WIENER_COEF_RANGE_UV[18][2]
  crates/rav2d/src/tables.rs:718   DEBLOCK_SIDE_THRESHOLDS[296]
  crates/rav2d/src/tables.rs:920   PC_WIENER_FILTERS[4][64][13]
  crates/rav2d/src/tables.rs:1187  PC_WIENER_SUB_CLASSIFY[4][256]
  crates/rav2d/src/tables.rs:1242  PC_WIENER_SUB_CLASSIFY_NS[4][7][256]
  crates/rav2d/src/tables.rs:2051  GAUSSIAN_SEQUENCE[2048]
  crates/rav2d/src/tables.rs:2172  PC_WIENER_LUT_TO_CLASS[4096]
  crates/rav2d/src/gdf_tables.rs:1    GDF_ALPHA[3168]
  crates/rav2d/src/gdf_tables.rs:162  GDF_WEIGHT[9504]
  crates/rav2d/src/gdf_tables.rs:548  GDF_BIAS[36][3]
  crates/rav2d/src/gdf_tables.rs:587  GDF_INTRA_ERROR[24576]
  crates/rav2d/src/gdf_tables.rs:1538 GDF_INTER_ERROR[30000]
  crates/rav2d/src/quantizer.rs:42    DQ_LOOKUP_TBL[24]
"""
import os
import re
import sys

TABLES = [
    # (file, rust name, c name, c type, shape)
    ("tables.rs", "CDEF_DIRECTIONS", "R2T_CDEF_DIRECTIONS", "int8_t", (12, 2)),
    ("tables.rs", "CCSO_QUANT_SZ", "R2T_CCSO_QUANT_SZ", "uint16_t", (4, 4)),
    ("tables.rs", "CCSO_OFFSET", "R2T_CCSO_OFFSET", "int8_t", (4, 8)),
    ("tables.rs", "NS_WIENER_COEF_RANGE_Y", "R2T_NS_WIENER_COEF_RANGE_Y", "int8_t", (16, 2)),
    ("tables.rs", "NS_WIENER_COEF_RANGE_UV", "R2T_NS_WIENER_COEF_RANGE_UV", "int8_t", (18, 2)),
    ("tables.rs", "DEBLOCK_SIDE_THRESHOLDS", "R2T_DEBLOCK_SIDE_THRESHOLDS", "int16_t", (296,)),
    ("tables.rs", "PC_WIENER_FILTERS", "R2T_PC_WIENER_FILTERS", "int16_t", (4, 64, 13)),
    ("tables.rs", "PC_WIENER_SUB_CLASSIFY", "R2T_PC_WIENER_SUB_CLASSIFY", "uint8_t", (4, 256)),
    ("tables.rs", "PC_WIENER_SUB_CLASSIFY_NS", "R2T_PC_WIENER_SUB_CLASSIFY_NS", "uint8_t", (4, 7, 256)),
    ("tables.rs", "GAUSSIAN_SEQUENCE", "R2T_GAUSSIAN_SEQUENCE", "int16_t", (2048,)),
    ("tables.rs", "PC_WIENER_LUT_TO_CLASS", "R2T_PC_WIENER_LUT_TO_CLASS", "uint8_t", (4096,)),
    ("gdf_tables.rs", "GDF_ALPHA", "R2T_GDF_ALPHA", "uint16_t", (3168,)),
    ("gdf_tables.rs", "GDF_WEIGHT", "R2T_GDF_WEIGHT", "int16_t", (9504,)),
    ("gdf_tables.rs", "GDF_BIAS", "R2T_GDF_BIAS", "int16_t", (36, 3)),
    ("gdf_tables.rs", "GDF_INTRA_ERROR", "R2T_GDF_INTRA_ERROR", "int8_t", (24576,)),
    ("gdf_tables.rs", "GDF_INTER_ERROR", "R2T_GDF_INTER_ERROR", "int8_t", (30000,)),
    ("quantizer.rs", "DQ_LOOKUP_TBL", "R2T_DQ_LOOKUP", "uint8_t", (24,)),
]

_DECL = r"(?:pub\s+)?(?:static|const)\s+{name}\s*:[^=]*=\s*\["


def _strip_comments(s):
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    return re.sub(r"//[^\n]*", " ", s)


def extract(src, name):
    m = re.search(_DECL.format(name=re.escape(name)), src)
    if not m:
        raise SystemExit("table %s not found" % name)
    i = m.end() - 1
    depth = 0
    j = i
    while True:
        c = src[j]
        if c == "[":
            depth += 1
        elif c == "]":
            depth -= 1
            if depth == 0:
                break
        j += 1
    body = _strip_comments(src[i:j + 1])
    return [int(v, 0) for v in re.findall(r"-?(?:0x[0-9a-fA-F]+|\d+)", body)]


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    base = os.path.join(root, "crates", "rav2d", "src")
    cache = {}
    out = [
        "/* Generated by extract_tables.py from the rav2d checkout; do not edit.",
        " * Table data of rav2d (BSD-2-Clause, Copyright (c) 2026 rav2d contributors;",
        " * parts Copyright (c) 2018-2026 VideoLAN and dav2d authors, Two Orioles LLC).",
        " * Redistribution of this data must retain that notice.  The values are the",
        " * reference's own: the filters are specified by them. */",
        "#pragma once",
        "#include <stdint.h>",
        "",
    ]
    for fname, rname, cname, ctype, shape in TABLES:
        if fname not in cache:
            with open(os.path.join(base, fname)) as f:
                cache[fname] = f.read()
        vals = extract(cache[fname], rname)
        n = 1
        for d in shape:
            n *= d
        if len(vals) != n:
            raise SystemExit("%s: %d values, expected %d" % (rname, len(vals), n))
        dims = "".join("[%d]" % d for d in shape)
        out.append("/* %s (%s) */" % (rname, fname))
        out.append("static const %s %s%s = {" % (ctype, cname, dims))
        # emit flat with braces elided (C allows brace elision for nested arrays)
        for k in range(0, n, 16):
            out.append("  " + ", ".join(str(v) for v in vals[k:k + 16]) + ",")
        out.append("};")
        out.append("")
    sys.stdout.write("\n".join(out))


if __name__ == "__main__":
    main()
