#!/usr/bin/env python3
"""Writes lr_taps.json: the tap-offset tables of rav2d's loop restoration
(crates/rav2d/src/looprestoration.rs:333-406: GDF_COORDS,
WIENER_NS_CONFIG_UV, WIENER_NS_CONFIG_UV_FROM_Y, PC_WIENER_CONFIG,
WIENER_NS_CONFIG_Y; each a list of (dy, dx) pairs) and the constants of the
PC-Wiener / multi-class classifier (looprestoration.rs:13-22:
PC_WIENER_NORMALIZER, MODE_WEIGHTS, MODE_OFFSETS; flat lists of integers),
read as data from the reference checkout.  Run once where the checkout exists:

    python tests/golden/extract_lr_taps.py /root/reference
"""
import json
import os
import re
import sys

NAMES = ["WIENER_NS_CONFIG_Y", "WIENER_NS_CONFIG_UV", "WIENER_NS_CONFIG_UV_FROM_Y", "PC_WIENER_CONFIG",
         "GDF_COORDS"]
# (name, element count)
CONSTANTS = [("PC_WIENER_NORMALIZER", 4), ("MODE_WEIGHTS", 12), ("MODE_OFFSETS", 4)]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, "crates/rav2d/src/looprestoration.rs")).read()
    out = {"source": "crates/rav2d/src/looprestoration.rs"}
    for name in NAMES:
        m = re.search(r"static %s: \[\[i8; 2\]; (\d+)\] = \[(.*?)\];" % name, text, re.S)
        if not m:
            raise SystemExit("table %s not found" % name)
        pairs = [[int(a), int(b)] for a, b in re.findall(r"\[\s*(-?\d+),\s*(-?\d+)\s*\]", m.group(2))]
        if len(pairs) != int(m.group(1)):
            raise SystemExit("table %s: %d entries, expected %s" % (name, len(pairs), m.group(1)))
        out[name] = pairs
    for name, count in CONSTANTS:
        m = re.search(r"(?:static|const) %s: [^=]*= \[(.*?)\];" % name, text, re.S)
        if not m:
            raise SystemExit("table %s not found" % name)
        vals = [int(v) for v in re.findall(r"-?\d+", m.group(1))]
        if len(vals) != count:
            raise SystemExit("table %s: %d entries, expected %d" % (name, len(vals), count))
        out[name] = vals
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lr_taps.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
