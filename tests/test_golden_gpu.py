"""The GPU path against the committed golden fixtures at the workloads' real sizes, bit for bit: the plain-LIF
workload (16,384 neurons, 1,000 steps) on Brain (per-step spike totals, raster and state sha256); 100 frames of 10 ms
(10,000 steps) of the full-scale fly (167,106 neurons, LIF + optic lobe) on FlyBrain with captured frames (per-frame
spike totals, final LIF and optic state sha256); and the 128 uneven subgraphs of the packed workload stepped 5,000
steps in one graph replay after a tuning pass, as the benchmark runs them (every instance's spike total and final
state sha256 against its solo oracle run)."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

from flyhip import synth
from flyhip.brain import Brain, LIFParams
from flyhip.fly import FlyBrain
from flyhip.optic import OpticParams

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def _mg():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", GOLD / "make_golden.py")
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


def test_plain_lif_full_size_matches_golden():
    mg = _mg()
    want = json.loads((GOLD / "c0_plain_lif.json").read_text())
    C0 = want["config"]
    c, inputs, hz = synth.plain_lif_graph(C0["graph_seed"], n=C0["n"], n_inputs=C0["n_inputs"], hz=C0["hz"])
    brain = Brain(c, LIFParams.shiu(dt=0.1), device="cuda", batch=1, seed=C0["seed"])
    brain.record_activity = True
    brain.set_poisson(inputs, hz)
    rows, totals = [], []
    for _ in range(C0["steps"]):
        brain.step()
        s = brain.spikes[0].cpu().numpy()
        rows.append(s)
        totals.append(int(s.sum()))
    assert totals == want["spike_totals"]
    assert mg.raster_digest(rows) == want["raster_sha256"]
    st = [getattr(brain, k)[0].cpu().numpy() for k in ("v", "g", "refrac", "rate", "spike_counts")]
    assert mg.digest(*st) == want["state_sha256"]


def test_full_scale_fly_frames_match_golden():
    mg = _mg()
    want = json.loads((GOLD / "c1_fly_frames.json").read_text())
    C1 = want["config"]
    conn, retina, layout = synth.fly_graph(C1["graph_seed"], C1["scale"])
    lp = LIFParams.shiu(dt=0.1)
    op = OpticParams(tau_by_type=conn.rate_tau, baseline_by_type=conn.rate_baseline, pair_gain=[])
    idx, hz = synth.fly_inputs(C1["input_seed"], layout, 300, 200.0)
    fb = FlyBrain(conn, retina, batch=1, device="cuda", seed=C1["seed"], lif_params=lp, optic_params=op,
                  cuda_graphs=True)
    fb.set_poisson(idx, hz)
    fb.random_movie(C1["movie_seed"])
    before = 0.0
    for frame in range(C1["frames"]):
        fb.step(10.0)                                     # one frame = frame_steps (100) steps of 0.1 ms
        total = float(fb.brain.spike_counts.sum().item())
        assert int(total - before) == want["frame_totals"][frame], f"frame {frame}"
        before = total
    b, o = fb.brain, fb.optic
    st = [getattr(b, k)[0].cpu().numpy() for k in ("v", "g", "refrac", "rate", "spike_counts")]
    assert mg.digest(*st) == want["state_sha256"]
    pr_cell = torch.as_tensor(retina.cells()["pr_cell"].astype(np.int64), device=o.I_lp.device)
    ost = [t[0].cpu().numpy() for t in (o.v, o.adapt, o.I_lp[:, pr_cell], o.I_mean[:, pr_cell])]
    assert mg.digest(*ost) == want["optic_sha256"]


def test_packed_workload_matches_golden_as_the_benchmark_runs_it():
    from flyhip.autotune import tune_packed
    from flyhip.packed import PackedBrains
    mg = _mg()
    want = json.loads((GOLD / "c3_packed.json").read_text())
    C3 = want["config"]
    mats, inputs, seeds = mg.c3_workload()
    lp = LIFParams.shiu(dt=0.1)
    pb = PackedBrains(mats, lp, seeds=seeds, inputs=inputs, device="cuda")
    del mats
    tune_packed(pb, use_cache=False)                  # leaves the fastest shape configured, back at step 0
    pb.run(C3["steps"])
    torch.cuda.synchronize()
    pb.check()
    bad = []
    for k, ref in want["instances"].items():
        st = pb.instance_state(int(k))
        total = int(st["counts"].sum())
        sha = mg.digest(st["v"], st["g"], st["refrac"], st["rate"], st["counts"])
        if total != ref["count_totals"][-1] or sha != ref["state_sha256"]:
            bad.append(k)
    assert not bad, f"{len(bad)} of {len(want['instances'])} instances differ, e.g. {bad[:5]}"
