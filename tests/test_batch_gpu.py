"""The batched engine against the CPU oracle: every lane is one fly with its own seeds, movie, Poisson cells and
rollout length, and must equal that fly's solo run on the oracle bit for bit (spikes every step, LIF and optic
state), including lanes that are refilled with a new rollout mid-run."""
import numpy as np
import pytest
import torch

from flyhip import synth
from flyhip.batch import BatchBrain, Rollout
from helpers import fly_setup, grouped_retina, oracle_fly

pytestmark = pytest.mark.gpu
DEV = "cuda"


def solo_oracle(conn, retina, lp, op, r: Rollout):
    om = oracle_fly(conn, retina, lp, op, 1, [r.seed], [r.movie_seed])
    om.st.poisson_p[0, r.input_idx] = (r.input_hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    return om


def rollouts(layout, n, frames_lo, frames_hi, frame_steps, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        idx, hz = synth.fly_inputs(1000 + k, layout, 30)
        out.append(Rollout(seed=int(rng.integers(1, 2**62)), movie_seed=int(rng.integers(1, 2**62)),
                           steps=int(rng.integers(frames_lo, frames_hi + 1)) * frame_steps, input_idx=idx, input_hz=hz))
    return out


@pytest.mark.parametrize("graph,mode,grouped", [(False, 0, False), (True, 1, False), (True, 2, False), (False, 2, False),
                                                (True, 3, False), (False, 3, False), (True, 4, False),
                                                (False, 4, False), (True, 5, False), (False, 5, False),
                                                (True, 5, True), (True, 6, False), (False, 6, True)])
def test_lanes_equal_solo_oracle_runs_with_refill(graph, mode, grouped):
    conn, retina, layout, lp, op = fly_setup(0.02, seed=2)
    if grouped:
        retina = grouped_retina(retina)
    frame_ms, frame_steps = 1.0, 10           # 10 LIF steps and one optic substep per frame: the ring holds a frame
    bb = BatchBrain(conn, retina, lp, op, device=DEV, frame_ms=frame_ms, grid=64, fold_mode=mode)
    rs = rollouts(layout, 70, 3, 5, frame_steps, seed=1)    # 70 rollouts > 64 lanes: six lanes are refilled
    rs[0].steps = 2 * frame_steps                             # lane 0 finishes first and takes rollout 64
    bb.submit(rs)
    bb.prepare(graph)
    lanes = [0, 5, 63]
    tracked, solos = {}, {}
    for frame in range(8):
        bb._refill()
        running = {l: bb.lane_rollout[l] for l in lanes if bb.lane_left[l] > 0}
        T0 = frame * frame_steps
        bb.run(1)
        torch.cuda.synchronize()
        raster = bb.ring_raster(T0, T0 + frame_steps)
        for l, r in running.items():
            if tracked.get(l) is not r:
                tracked[l], solos[l] = r, solo_oracle(conn, retina, lp, op, r)
            om = solos[l]
            steps = []
            om.frame(frame_steps, steps)
            for k in range(frame_steps):
                assert np.array_equal(raster[k, l], steps[k][0]), f"lane {l} frame {frame} step {k}"
            st = bb.lane_state(l)
            for key in ("v", "g", "refrac", "rate", "spike_counts"):
                assert np.array_equal(st[key], getattr(om.st, key)[0]), f"{key} lane {l} frame {frame}"
            assert np.array_equal(bb.optic_state(l)["v"], om.os.v[0]), f"optic v lane {l} frame {frame}"
    assert tracked[0] is rs[64], "lane 0 should have been refilled with the first queued rollout"
    assert bb.completed > 0
    assert bb.counts.sum().item() > 0


def test_full_scale_batch_is_deterministic_and_matches_plain_path():
    """Full-scale graph, 64 flies, 3 frames of 10 ms: three runs (fold modes 2, 4 and 5) give the same bits, and lane 7
    equals the plain single-fly FlyBrain (batch=1, same seeds and inputs) bit for bit."""
    from flyhip.brain import LIFParams
    from flyhip.fly import FlyBrain
    from flyhip.optic import OpticParams
    conn, retina, layout = synth.fly_graph(1, 1.0)
    lp = LIFParams.shiu(dt=0.1)
    op = OpticParams(tau_by_type=conn.rate_tau, baseline_by_type=conn.rate_baseline, pair_gain=[])
    rs = rollouts(layout, 64, 3, 3, 100, seed=5)
    states = []
    for rep, mode in enumerate((2, 4, 5, 6)):
        bb = BatchBrain(conn, retina, lp, op, device=DEV, fold_mode=mode)
        bb.submit(rs)
        bb.run(3)
        torch.cuda.synchronize()
        states.append({k: getattr(bb, k).cpu().clone() for k in ("v", "g", "refrac", "rate", "counts", "ov", "ring")})
        if rep == 0:
            lane7 = bb.lane_state(7)
            optic7 = bb.optic_state(7)
            total = float(bb.counts.sum())
        del bb
    for other in states[1:]:
        for k in states[0]:
            assert torch.equal(states[0][k], other[k]), k
    assert total > 0
    r = rs[7]
    fb = FlyBrain(conn, retina, batch=1, device=DEV, seed=r.seed, lif_params=lp, optic_params=op)
    fb.set_poisson(r.input_idx, r.input_hz)
    fb.random_movie(r.movie_seed)
    for _ in range(3):
        fb.step(10.0)
    for key, t in (("v", fb.brain.v), ("g", fb.brain.g), ("rate", fb.brain.rate), ("refrac", fb.brain.refrac),
                   ("spike_counts", fb.brain.spike_counts)):
        assert np.array_equal(t[0].cpu().numpy(), lane7[key]), key
    assert np.array_equal(fb.optic.v[0].cpu().numpy(), optic7["v"])


def test_full_scale_uneven_rollouts_with_refill_equal_solo_runs():
    """The benchmark's B=64 workload at full scale (rollouts of 2,000-10,000 steps, 300 Poisson inputs each, one
    captured graph per frame, fold mode 6 with 6 steps per launch and 128 workgroups per CU): lane 0's rollout is
    shortened to 2,000 steps, so after 20 frames the lane is reset and takes the first queued rollout.  Lane 0 (its
    first rollout at frame 20, the refilled one at frame 25) and two lanes whose rollouts run throughout equal the
    plain FlyBrain (batch=1) run of their rollout, bit for bit."""
    import bench
    from flyhip.brain import LIFParams
    from flyhip.fly import FlyBrain
    from flyhip.optic import OpticParams
    conn, retina, layout = synth.fly_graph(1, 1.0)
    lp = LIFParams.shiu(dt=0.1)
    op = OpticParams(tau_by_type=conn.rate_tau, baseline_by_type=conn.rate_baseline, pair_gain=[])
    rs = bench.make_rollouts(layout, 70, seed=11)
    rs[0].steps = 2000
    bb = BatchBrain(conn, retina, lp, op, device=DEV)
    bb.submit(rs)
    bb.prepare(graph=True)
    bb.configure(fold_mode=6, steps_per_launch=6, grid_per_cu=128)

    def solo(r, frames):
        fb = FlyBrain(conn, retina, batch=1, device=DEV, seed=r.seed, lif_params=lp, optic_params=op, cuda_graphs=True)
        fb.set_poisson(r.input_idx, r.input_hz)
        fb.random_movie(r.movie_seed)
        for _ in range(frames):
            fb.step(10.0)
        torch.cuda.synchronize()
        out = {k: getattr(fb.brain, k)[0].cpu().numpy() for k in ("v", "g", "refrac", "rate", "spike_counts")}
        out["optic_v"] = fb.optic.v[0].cpu().numpy()
        del fb
        return out

    def same(lane_state, optic_state, ref, what):
        for k in ("v", "g", "refrac", "rate", "spike_counts"):
            assert np.array_equal(lane_state[k], ref[k]), f"{what}: {k}"
        assert np.array_equal(optic_state["v"], ref["optic_v"]), f"{what}: optic v"

    bb.run(20)
    torch.cuda.synchronize()
    assert bb.lane_rollout[0] is rs[0] and bb.lane_left[0] == 0
    same(bb.lane_state(0), bb.optic_state(0), solo(rs[0], 20), "lane 0, first rollout")
    bb.run(5)
    torch.cuda.synchronize()
    assert bb.lane_rollout[0] is rs[64] and bb.h_frame[0] == 5
    same(bb.lane_state(0), bb.optic_state(0), solo(rs[64], 5), "lane 0, refilled")
    steady = [l for l in range(1, 64) if rs[l].steps >= 2500][:2]
    assert len(steady) == 2
    for l in steady:
        assert bb.lane_rollout[l] is rs[l]
        same(bb.lane_state(l), bb.optic_state(l), solo(rs[l], 25), f"lane {l}")
    assert bb.completed >= 1
