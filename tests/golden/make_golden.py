"""Writes the golden fixtures of this directory from the CPU oracle (oracle/):
  * c0_plain_lif.json: the plain-LIF workload at its full size (16,384 neurons, 200 Poisson inputs at 150 Hz, 1,000
    steps of dt 0.1 ms): the spike total of every step, sha256 of the packed raster and of the final state;
  * c1_fly_frames.json: the full-scale fly (167,106 neurons with the optic lobe, one fly) for 100 frames of 10 ms,
    i.e. 10,000 steps: the spike total of every step and of every frame, sha256 of the raster, of the final LIF
    state and of the final optic state;
  * c3_packed.json: the 128 uneven subgraphs of the packed workload (flyhip.synth.packed_workload), each run alone
    for 5,000 steps: per instance the spike total every 500 steps and sha256 of its final state.
A run of the GPU path at the same seeds is compared with them without re-running the oracle.

    python tests/golden/make_golden.py            # rewrites all three (C3 uses a pool of processes, a few minutes)
    python tests/golden/make_golden.py c1         # only the named ones
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
for p in (ROOT, ROOT / "flyverse-core_amd", ROOT / "tests"):
    sys.path.insert(0, str(p))

C0 = dict(graph_seed=0, n=16_384, n_inputs=200, hz=150.0, seed=20240607, steps=1000)
C1 = dict(graph_seed=1, scale=1.0, input_seed=77, seed=987654321, movie_seed=13579, frames=100, frame_steps=100)
C3 = dict(seed=3, count=128, n_min=5_000, n_max=80_000, mean_in_degree=153.0, input_seed=12, steps=5000, every=500)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def raster_digest(rows) -> str:
    return digest(*(np.packbits(r.astype(bool).reshape(-1)) for r in rows))


def c0_oracle():
    from flyhip import synth
    from flyhip.brain import LIFParams
    from helpers import oracle_fly
    c, inputs, hz = synth.plain_lif_graph(C0["graph_seed"], n=C0["n"], n_inputs=C0["n_inputs"], hz=C0["hz"])
    lp = LIFParams.shiu(dt=0.1)
    om = oracle_fly(c, None, lp, None, 1, [C0["seed"]], None, with_optic=False)
    om.st.poisson_p[0, inputs] = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    rows, totals = [], []
    for st in om.lif_steps(C0["steps"]):
        rows.append(st.spikes[0].copy())
        totals.append(int(st.spikes[0].sum()))
    s = om.st
    return {"config": C0, "spike_totals": totals, "raster_sha256": raster_digest(rows),
            "state_sha256": digest(s.v, s.g, s.refrac, s.rate, s.spike_counts)}


class RasterHash:
    """A list-like sink for the oracle's per-step spike rows: keeps each row's total and hashes the packed rows in
    order (the same digest as raster_digest over the list), without holding 10,000 rows of 167,106 floats."""

    def __init__(self):
        self.h = hashlib.sha256()
        self.totals = []

    def append(self, row):
        self.h.update(np.ascontiguousarray(np.packbits(row.astype(bool).reshape(-1))).tobytes())
        self.totals.append(int(row.sum()))


def c1_oracle():
    from flyhip import synth
    from flyhip.brain import LIFParams
    from flyhip.optic import OpticParams
    from helpers import oracle_fly
    conn, retina, layout = synth.fly_graph(C1["graph_seed"], C1["scale"])
    lp = LIFParams.shiu(dt=0.1)
    op = OpticParams(tau_by_type=conn.rate_tau, baseline_by_type=conn.rate_baseline, pair_gain=[])
    idx, hz = synth.fly_inputs(C1["input_seed"], layout, 300, 200.0)
    om = oracle_fly(conn, retina, lp, op, 1, [C1["seed"]], [C1["movie_seed"]])
    om.st.poisson_p[0, idx] = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    om.poisson_cells = np.asarray(idx, np.int32)
    rows = RasterHash()
    for _ in range(C1["frames"]):
        om.frame(C1["frame_steps"], rows)
    s, o = om.st, om.os
    totals = rows.totals
    fs = C1["frame_steps"]
    return {"config": C1, "spike_totals": totals, "raster_sha256": rows.h.hexdigest(),
            "frame_totals": [sum(totals[k * fs:(k + 1) * fs]) for k in range(C1["frames"])],
            "state_sha256": digest(s.v, s.g, s.refrac, s.rate, s.spike_counts),
            # the photoreceptor state as seen by each photoreceptor (its (column, family) cell's values)
            "optic_sha256": digest(o.v, o.adapt, *(x[:, retina.cells()["pr_cell"]] for x in (o.I_lp, o.I_mean)))}


def c3_workload(which=None):
    from flyhip import synth
    return synth.packed_workload(C3["seed"], C3["count"], C3["n_min"], C3["n_max"], C3["mean_in_degree"],
                                 C3["input_seed"], which=which)


def c3_instance(args):
    """Instance k of the packed workload run alone on the oracle: spike totals every C3["every"] steps and the final
    state's sha256 (v, g, refrac, rate, counts)."""
    m, (idx, hz), seed = args
    from flyhip.brain import LIFParams
    from flyhip.packed import _instance_connectome
    from helpers import oracle_fly
    lp = LIFParams.shiu(dt=0.1)
    om = oracle_fly(_instance_connectome(m), None, lp, None, 1, [int(seed)], None, with_optic=False)
    om.st.poisson_p[0, idx] = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    om.poisson_cells = np.asarray(idx, np.int32)
    totals = []
    for k in range(C3["steps"] // C3["every"]):
        for _ in om.lif_steps(C3["every"]):
            pass
        totals.append(int(om.st.spike_counts.sum()))
    s = om.st
    return {"cells": int(m.shape[0]), "edges": int(m.nnz), "count_totals": totals,
            "state_sha256": digest(s.v[0], s.g[0], s.refrac[0], s.rate[0], s.spike_counts[0])}


def c3_oracle(instances=None, workers=8):
    ks = list(range(C3["count"]) if instances is None else instances)
    mats, inputs, seeds = c3_workload(ks)
    jobs = list(zip(mats, inputs, seeds))
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            results = pool.map(c3_instance, jobs, chunksize=1)
    else:
        results = [c3_instance(j) for j in jobs]
    return {"config": C3, "instances": {str(k): r for k, r in zip(ks, results)}}


FIXTURES = {"c0": ("c0_plain_lif.json", c0_oracle), "c1": ("c1_fly_frames.json", c1_oracle),
            "c3": ("c3_packed.json", c3_oracle)}

if __name__ == "__main__":
    for key in sys.argv[1:] or list(FIXTURES):
        name, fn = FIXTURES[key]
        out = fn()
        (HERE / name).write_text(json.dumps(out, indent=0) + "\n")
        print(name, "written")
