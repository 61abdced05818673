#!/usr/bin/env python3
"""Writes table_digests.json: for every lookup table the filter stage takes
from the reference (the tables rav2d_amd/csrc/tables.h holds), its length and
the SHA-256 of its values as little-endian int64, read here from the
reference's Rust sources by a parser of its own (bracket matching over the
array literal, comments stripped), so that a test can check tables.h against
the reference without the checkout.  Run once where the checkout exists:

    python tests/golden/table_digests.py /root/reference
"""
import hashlib
import json
import os
import re
import sys

# (source file under crates/rav2d/src, Rust name, name in tables.h)
TABLES = [
    ("tables.rs", "CDEF_DIRECTIONS", "R2T_CDEF_DIRECTIONS"),
    ("tables.rs", "CCSO_QUANT_SZ", "R2T_CCSO_QUANT_SZ"),
    ("tables.rs", "CCSO_OFFSET", "R2T_CCSO_OFFSET"),
    ("tables.rs", "NS_WIENER_COEF_RANGE_Y", "R2T_NS_WIENER_COEF_RANGE_Y"),
    ("tables.rs", "NS_WIENER_COEF_RANGE_UV", "R2T_NS_WIENER_COEF_RANGE_UV"),
    ("tables.rs", "DEBLOCK_SIDE_THRESHOLDS", "R2T_DEBLOCK_SIDE_THRESHOLDS"),
    ("tables.rs", "PC_WIENER_FILTERS", "R2T_PC_WIENER_FILTERS"),
    ("tables.rs", "PC_WIENER_SUB_CLASSIFY", "R2T_PC_WIENER_SUB_CLASSIFY"),
    ("tables.rs", "PC_WIENER_SUB_CLASSIFY_NS", "R2T_PC_WIENER_SUB_CLASSIFY_NS"),
    ("tables.rs", "GAUSSIAN_SEQUENCE", "R2T_GAUSSIAN_SEQUENCE"),
    ("tables.rs", "PC_WIENER_LUT_TO_CLASS", "R2T_PC_WIENER_LUT_TO_CLASS"),
    ("gdf_tables.rs", "GDF_ALPHA", "R2T_GDF_ALPHA"),
    ("gdf_tables.rs", "GDF_WEIGHT", "R2T_GDF_WEIGHT"),
    ("gdf_tables.rs", "GDF_BIAS", "R2T_GDF_BIAS"),
    ("gdf_tables.rs", "GDF_INTRA_ERROR", "R2T_GDF_INTRA_ERROR"),
    ("gdf_tables.rs", "GDF_INTER_ERROR", "R2T_GDF_INTER_ERROR"),
    ("quantizer.rs", "DQ_LOOKUP_TBL", "R2T_DQ_LOOKUP"),
]

_INT = re.compile(r"(-?)\s*(0x[0-9a-fA-F_]+|[0-9][0-9_]*)(?:[iu](?:8|16|32|64|size))?")


def literal_values(text, name):
    m = re.search(r"\b%s\s*:" % re.escape(name), text)
    if not m:
        raise SystemExit("%s not found" % name)
    start = text.index("=", m.end())
    start = text.index("[", start)
    depth, i = 0, start
    while True:
        c = text[i]
        if c == "[":
            depth += 1
        elif c == "]":
            depth -= 1
            if depth == 0:
                break
        i += 1
    body = re.sub(r"/\*.*?\*/", " ", text[start:i + 1], flags=re.S)
    body = re.sub(r"//[^\n]*", " ", body)
    out = []
    for sign, digits in _INT.findall(body):
        v = int(digits.replace("_", ""), 0)
        out.append(-v if sign else v)
    return out


def digest(values):
    b = b"".join(int(v).to_bytes(8, "little", signed=True) for v in values)
    return hashlib.sha256(b).hexdigest()


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = {}
    for fname, rust, cname in TABLES:
        text = open(os.path.join(ref, "crates/rav2d/src", fname)).read()
        vals = literal_values(text, rust)
        out[cname] = {"source": "crates/rav2d/src/%s %s" % (fname, rust), "n": len(vals), "sha256": digest(vals)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "table_digests.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
