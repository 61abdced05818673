"""Writes chain_hashes.json: for a fixed set of seeded synthetic frames, the
64-bit plane hashes (rav2d_amd.plane_checksum) of the CPU oracle's output
after each stage of the chain and of its film-grain display copy.

    python tests/golden/make_golden.py

The fixture freezes the oracle's results, so later changes to the oracle or
to the synthetic generator show up as a fixture mismatch rather than silently
moving both sides of the GPU parity tests."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import rav2d_amd as R  # noqa: E402
import oracle_lib as O  # noqa: E402

# (w, h, bitdepth, sb128, seed, tools, lossless_pct)
FRAMES = [
    (352, 288, 8, 0, 1, R.TOOL_ALL & ~R.TOOL_GRAIN, 0),   # CIF 8-bit, grain off
    (480, 272, 10, 1, 2, R.TOOL_ALL, 0),                  # 10-bit, sb128, every tool
    (136, 72, 10, 0, 3, R.TOOL_ALL, 0),                   # ragged superblocks
    (256, 192, 8, 1, 4, R.TOOL_ALL, 10),                  # some lossless blocks
    (64, 64, 10, 0, 5, R.TOOL_DEBLOCK | R.TOOL_GRAIN, 0),
    (128, 72, 8, 0, 6, R.TOOL_WIENER | R.TOOL_GDF, 0),
]

STAGES = [
    ("deblock_v", R.INLOOP_DEBLOCK | (1 << 8)),
    ("deblock", R.INLOOP_DEBLOCK),
    ("cdef_ccso", R.INLOOP_DEBLOCK | R.INLOOP_CDEF | R.INLOOP_CCSO),
    ("in_loop", R.INLOOP_ALL),
]


def compute():
    out = []
    for (w, h, bd, sb, seed, tools, ll) in FRAMES:
        sf = R.SynthFrame(w, h, bd, sb, seed, tools=tools, lossless_pct=ll)
        rec = {"frame": [w, h, bd, sb, seed, tools, ll], "input": R.plane_checksum(sf.planes())}
        for name, mask in STAGES:
            rec[name] = R.plane_checksum(O.filter_frame(sf, mask))
        rec["display"] = R.plane_checksum(O.apply_grain(sf, O.filter_frame(sf)))
        out.append(rec)
        sf.close()
    return out


if __name__ == "__main__":
    data = compute()
    with open(os.path.join(HERE, "chain_hashes.json"), "w") as f:
        json.dump({"hash": "sum of splitmix64(plane << 60 | y << 40 | x << 20 | value) mod 2^64",
                   "frames": data}, f, indent=1)
    print("wrote %d frames" % len(data))
