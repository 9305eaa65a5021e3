"""Writes the golden fixtures of this directory from the CPU oracle (oracle/): the plain-LIF workload at its full
size (16,384 neurons, 200 Poisson inputs at 150 Hz, 1,000 steps of dt 0.1 ms) and the full-scale fly (167,106
neurons with the optic lobe, one fly, 10 frames of 10 ms).  Each fixture holds the spike total of every step, the
sha256 of the packed raster and of the final state, so a run of the GPU path at the same seeds can be compared with
it without re-running the oracle.

    python tests/golden/make_golden.py            # rewrites c0_plain_lif.json and c1_fly_frames.json
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
C1 = dict(graph_seed=1, scale=1.0, input_seed=77, seed=987654321, movie_seed=13579, frames=10, frame_steps=100)


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
    rows = []
    for _ in range(C1["frames"]):
        om.frame(C1["frame_steps"], rows)
    s, o = om.st, om.os
    totals = [int(r.sum()) for r in rows]
    fs = C1["frame_steps"]
    return {"config": C1, "spike_totals": totals, "raster_sha256": raster_digest(rows),
            "frame_totals": [sum(totals[k * fs:(k + 1) * fs]) for k in range(C1["frames"])],
            "state_sha256": digest(s.v, s.g, s.refrac, s.rate, s.spike_counts),
            # the photoreceptor state as seen by each photoreceptor (its (column, family) cell's values)
            "optic_sha256": digest(o.v, o.adapt, *(x[:, retina.cells()["pr_cell"]] for x in (o.I_lp, o.I_mean)))}


if __name__ == "__main__":
    for name, fn in (("c0_plain_lif.json", c0_oracle), ("c1_fly_frames.json", c1_oracle)):
        out = fn()
        (HERE / name).write_text(json.dumps(out, indent=0) + "\n")
        print(name, sum(out["spike_totals"]), "spikes")
