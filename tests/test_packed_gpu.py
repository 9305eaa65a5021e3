"""Packed uneven subgraphs in one persistent launch against the CPU oracle: every instance equals its own
one-fly oracle run (spikes every step, state), across launches that continue one another."""
import numpy as np
import pytest
import torch

from flyhip import synth
from flyhip.brain import LIFParams
from flyhip.packed import PackedBrains, _instance_connectome
from helpers import oracle_fly

pytestmark = pytest.mark.gpu


def instances(seed=0, count=6):
    rng = np.random.default_rng(seed)
    mats = synth.packed_subgraphs(seed, count, 300, 2500, mean_in_degree=60.0)
    inputs = []
    for m in mats:
        n = m.shape[0]
        k = max(5, n // 25)
        inputs.append((np.sort(rng.choice(n, k, replace=False)), rng.uniform(50, 300, k).astype(np.float32)))
    seeds = rng.integers(1, 2**62, count, dtype=np.uint64)
    return mats, inputs, seeds


@pytest.mark.parametrize("window,mode", [(18, "graph"), (7, "eager"), (18, "coop"), (5, "coop")])
def test_instances_equal_solo_oracle_runs(window, mode):
    mats, inputs, seeds = instances()
    lp = LIFParams.shiu(dt=0.1)
    pb = PackedBrains(mats, lp, seeds=seeds, inputs=inputs, device="cuda", window=window, grid=64, mode=mode)
    solos = []
    for k, m in enumerate(mats):
        c = _instance_connectome(m)
        om = oracle_fly(c, None, lp, None, 1, [int(seeds[k])], None, with_optic=False)
        om.st.poisson_p[0, inputs[k][0]] = (inputs[k][1] * np.float32(lp.dt / 1000.0)).astype(np.float32)
        solos.append(om)
    chunk = 40
    for c0 in range(0, 200, chunk):
        pb.run(chunk)
        torch.cuda.synchronize()
        pb.check()
        raster = pb.ring_raster(c0, c0 + chunk)
        for k, om in enumerate(solos):
            s, e = pb.offsets[k], pb.offsets[k + 1]
            for j, st in enumerate(om.lif_steps(chunk)):
                assert np.array_equal(raster[j, s:e], st.spikes[0]), f"instance {k} step {c0 + j}"
            got = pb.instance_state(k)
            for key, ref in (("v", om.st.v[0]), ("g", om.st.g[0]), ("rate", om.st.rate[0]),
                             ("counts", om.st.spike_counts[0])):
                assert np.array_equal(got[key], ref), f"{key} instance {k} after {c0 + chunk} steps"
    assert pb.counts.sum().item() > 0


def test_full_size_instances_and_launch_shapes():
    """16 subgraphs of 5k-80k cells at the full graph's in-degree: the smallest, largest and a middle instance equal
    their solo oracle runs for 120 steps, and every window / grid gives the same trajectories."""
    count = 16
    mats = synth.packed_subgraphs(5, count, 5_000, 80_000, mean_in_degree=152.0)
    rng = np.random.default_rng(5)
    inputs = [(np.sort(rng.choice(m.shape[0], 300, replace=False)), rng.uniform(0, 200, 300).astype(np.float32))
              for m in mats]
    seeds = rng.integers(1, 2**62, count, dtype=np.uint64)
    lp = LIFParams.shiu(dt=0.1)
    pb = PackedBrains(mats, lp, seeds=seeds, inputs=inputs, device="cuda")
    sizes = [m.shape[0] for m in mats]
    check = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argsort(sizes)[count // 2])})
    solos = {}
    for k in check:
        om = oracle_fly(_instance_connectome(mats[k]), None, lp, None, 1, [int(seeds[k])], None, with_optic=False)
        om.st.poisson_p[0, inputs[k][0]] = (inputs[k][1] * np.float32(lp.dt / 1000.0)).astype(np.float32)
        solos[k] = om
    rasters = {k: [] for k in check}
    for c0 in range(0, 120, 40):
        pb.run(40)
        torch.cuda.synchronize()
        pb.check()
        raster = pb.ring_raster(c0, c0 + 40)
        for k, om in solos.items():
            s, e = pb.offsets[k], pb.offsets[k + 1]
            for j, st in enumerate(om.lif_steps(40)):
                assert np.array_equal(raster[j, s:e], st.spikes[0]), f"instance {k} step {c0 + j}"
            rasters[k].append(raster[:, s:e])
    for k, om in solos.items():
        got = pb.instance_state(k)
        assert np.array_equal(got["v"], om.st.v[0]) and np.array_equal(got["counts"], om.st.spike_counts[0])
    ref = {name: x.clone() for name, x in (("v", pb.v), ("counts", pb.counts), ("rate", pb.rate))}
    assert ref["counts"].sum().item() > 0
    for window, grid, mode in ((6, 256, "eager"), (9, pb.max_grid, "graph"), (18, 4096, "graph"), (18, 64, "coop")):
        pb.configure(window=window, grid=grid, mode=mode)
        for _ in range(3):
            pb.run(40)
        torch.cuda.synchronize()
        pb.check()
        for name, x in ref.items():
            assert torch.equal(getattr(pb, name), x), f"{name} differs with window {window} grid {grid} {mode}"
