"""The lookup tables of rav2d_amd/csrc/tables.h as numpy arrays (flattened, in
declaration order), parsed from the header's text."""
import os
import re

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "rav2d_amd", "csrc", "tables.h")
_DECL = re.compile(r"static const (\w+) (R2T_\w+)((?:\[\d+\])+) = \{(.*?)\};", re.S)
_cache = {}


def tables():
    if not _cache:
        text = re.sub(r"/\*.*?\*/", " ", open(PATH).read(), flags=re.S)
        for ctype, name, dims, body in _DECL.findall(text):
            vals = [int(v, 0) for v in re.findall(r"-?(?:0x[0-9a-fA-F]+|\d+)", body)]
            n = int(np.prod([int(d) for d in re.findall(r"\[(\d+)\]", dims)]))
            assert len(vals) == n, (name, len(vals), n)
            _cache[name] = np.array(vals, dtype=np.int64)
    return _cache
