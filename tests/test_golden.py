"""The oracle reproduces the committed golden fixtures (tests/golden/make_golden.py): the plain-LIF workload at its
full size and 100 frames (10,000 steps) of the full-scale fly, spike totals per step and sha256 of raster and final
state; and the two smallest of the 128 packed-workload instances over their 5,000 steps."""
import json
from pathlib import Path

import pytest

GOLD = Path(__file__).resolve().parent / "golden"


def _mg():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", GOLD / "make_golden.py")
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


@pytest.mark.parametrize("name,fn", [("c0_plain_lif.json", "c0_oracle"), ("c1_fly_frames.json", "c1_oracle")])
def test_oracle_reproduces_golden(name, fn):
    mg = _mg()
    want = json.loads((GOLD / name).read_text())
    got = getattr(mg, fn)()
    assert got["spike_totals"] == want["spike_totals"]
    for k in want:
        if k.endswith("sha256"):
            assert got[k] == want[k], k


def test_oracle_reproduces_the_smallest_packed_instances():
    mg = _mg()
    want = json.loads((GOLD / "c3_packed.json").read_text())
    assert want["config"] == mg.C3 and len(want["instances"]) == mg.C3["count"]
    smallest = sorted(want["instances"], key=lambda k: want["instances"][k]["cells"])[:2]
    got = mg.c3_oracle([int(k) for k in smallest], workers=1)["instances"]
    for k in smallest:
        assert got[k] == want["instances"][k], k
