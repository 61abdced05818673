"""Committed golden hashes (tests/golden/chain_hashes.json, written by
tests/golden/make_golden.py): the CPU oracle must still produce them, and the
HIP path must produce them stage by stage."""
import json
import os

import pytest

import rav2d_amd as R
import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = json.load(open(os.path.join(HERE, "golden", "chain_hashes.json")))["frames"]
STAGE_MASK = {
    "deblock_v": R.INLOOP_DEBLOCK | (1 << 8),
    "deblock": R.INLOOP_DEBLOCK,
    "cdef_ccso": R.INLOOP_DEBLOCK | R.INLOOP_CDEF | R.INLOOP_CCSO,
    "in_loop": R.INLOOP_ALL,
}
STAGE_PASS = {"deblock_v": R.PASS_DEBLOCK_V, "deblock": R.PASS_DEBLOCK_H, "cdef_ccso": R.PASS_CDEF}


def _synth(rec):
    w, h, bd, sb, seed, tools, ll = rec["frame"]
    return R.SynthFrame(w, h, bd, sb, seed, tools=tools, lossless_pct=ll)


@pytest.mark.parametrize("rec", DATA, ids=lambda r: "%dx%d_%db" % tuple(r["frame"][:3]))
def test_oracle_reproduces_golden(rec):
    sf = _synth(rec)
    try:
        assert R.plane_checksum(sf.planes()) == rec["input"], "synthetic generator changed"
        for name, mask in STAGE_MASK.items():
            assert R.plane_checksum(O.filter_frame(sf, mask)) == rec[name], name
        assert R.plane_checksum(O.apply_grain(sf, O.filter_frame(sf))) == rec["display"]
    finally:
        sf.close()


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("rec", DATA, ids=lambda r: "%dx%d_%db" % tuple(r["frame"][:3]))
def test_gpu_reproduces_golden(ctx, rec):
    sf = _synth(rec)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        assert f.checksum(0) == rec["input"]
        for name, p in STAGE_PASS.items():
            f.run_pass(p)
            assert f.checksum(1) == rec[name], name
        f.filter()
        assert f.checksum(1) == rec["in_loop"]
        f.apply_grain()
        assert f.checksum(2) == rec["display"]
    finally:
        f.close()
        sf.close()
