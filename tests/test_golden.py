"""The oracle reproduces the committed golden fixtures (tests/golden/make_golden.py): the plain-LIF workload at its
full size and ten frames of the full-scale fly, spike totals per step and sha256 of raster and final state."""
import json
from pathlib import Path

import pytest

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name,fn", [("c0_plain_lif.json", "c0_oracle"), ("c1_fly_frames.json", "c1_oracle")])
def test_oracle_reproduces_golden(name, fn):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", GOLD / "make_golden.py")
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    want = json.loads((GOLD / name).read_text())
    got = getattr(mg, fn)()
    assert got["spike_totals"] == want["spike_totals"]
    for k in want:
        if k.endswith("sha256"):
            assert got[k] == want[k], k
