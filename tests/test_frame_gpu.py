"""The fused per-frame optic launches (csrc/optic.hip k_frame_cells + k_frame_held at frame start, k_frame_out at
frame end; flyverse/optic.py:496-517, 572, 579-584, 599-604) against the CPU oracle's frame (oracle.optic_frame), bit
for bit: frames with and without substeps, carried remainders, lanes reset mid-run (fresh photoreceptor state), and
a frame replayed from a captured graph."""
import numpy as np
import pytest
import torch

import oracle
from flyhip.optic import OpticLobe, optic_matrices
from helpers import fly_setup, grouped_retina

pytestmark = pytest.mark.gpu
DEV = "cuda"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def N(t):
    return t.detach().cpu().numpy()


def assert_bits(a, b, what=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    if bad.any():
        i = tuple(np.argwhere(bad)[0])
        raise AssertionError(f"{what}: {bad.sum()} of {a.size} differ, first at {i}: {a[i]!r} vs {b[i]!r}")


def _setup(B, seed=2):
    conn, retina, layout, lp, op = fly_setup(0.02, seed=seed)
    retina = grouped_retina(retina, per_column=3, seed=seed + 1)
    m = optic_matrices(conn, retina, op)
    spec = oracle.OptSpec(m, op, retina.pr_sens, conn.n, retina.cells())
    ol = OpticLobe(conn, retina, op, device=DEV, batch=B)
    os_ = oracle.OpticState(B, ol.n_rate, retina.n_cells)
    return conn, retina, ol, spec, os_


def _check_state(ol, os_, spec, what):
    assert_bits(N(ol.v), os_.v, f"v {what}")
    assert_bits(N(ol.adapt), os_.adapt, f"adapt {what}")
    assert_bits(N(ol.I_lp), os_.I_lp, f"I_lp {what}")
    assert_bits(N(ol.I_mean), os_.I_mean, f"I_mean {what}")
    assert_bits(N(ol.delta_rate), oracle.optic_dr(os_.v, spec.b_vec), f"delta_rate {what}")
    assert_bits(N(ol._fresh), os_.fresh, f"fresh {what}")


@pytest.mark.parametrize("substep", ["panel", "grouped"])
def test_fused_frames_match_the_oracle_frames(substep):
    """Twelve frames of B=3 flies: 10 ms (10 substeps), 0.4 ms (none; the remainder carries), 2.5 ms (2 or 3 with
    the remainder), with a fly reset after frame 4; drive, optic state, photoreceptor state and delta_rate bit-exact.
    Both substep kernels: the W_rr product by persistent panels (the default) and by row groups."""
    from flyhip import hip
    B = 3
    conn, retina, ol, spec, os_ = _setup(B)
    if substep != "panel":
        ol.W_rr = hip.CSR(ol.matrices["M_rr"], DEV, kernel=substep)
    assert ol.W_rr.kernel == substep
    rng = np.random.default_rng(11)
    lengths = [10.0, 0.4, 2.5, 10.0, 0.4, 0.4, 2.5, 10.0, 2.5, 0.4, 10.0, 10.0]
    substeps = 0
    for f, ms in enumerate(lengths):
        rad = rng.random((B, retina.n_col, 4)).astype(np.float32)
        rate = (rng.random((B, conn.n)) * 400.0 - 50.0).astype(np.float32)   # the clamp of s is exercised
        substeps += ol.frame_substeps(ms)
        out = torch.zeros(B, conn.n, device=DEV)
        got = ol.step_frame(T(rad), T(rate), ms, out=out)
        want = oracle.optic_frame(os_, spec, rad, rate, ms)
        assert_bits(N(got), want, f"drive frame {f} ({ms} ms)")
        _check_state(ol, os_, spec, f"frame {f}")
        if f == 4:
            ol.reset([1])
            os_.fresh[1] = 1
            os_.v[1] = 0.0
            os_.adapt[1] = 0.0
    assert substeps > 40
    assert np.abs(N(out)).max() > 0


def test_frame_replays_from_a_captured_graph():
    """One 10 ms frame captured in a graph and replayed four times, eager work and synchronisation between replays:
    every replay matches the oracle."""
    B = 2
    conn, retina, ol, spec, os_ = _setup(B, seed=5)
    rng = np.random.default_rng(3)
    rad_t = torch.zeros(B, retina.n_col, 4, device=DEV)
    rate_t = torch.zeros(B, conn.n, device=DEV)
    out = torch.zeros(B, conn.n, device=DEV)
    ol.prepare(10.0)
    # a first eager frame (fresh state), then the capture
    rad = rng.random((B, retina.n_col, 4)).astype(np.float32)
    rate = (rng.random((B, conn.n)) * 300.0).astype(np.float32)
    rad_t.copy_(T(rad))
    rate_t.copy_(T(rate))
    assert_bits(N(ol.step_frame(rad_t, rate_t, 10.0, out=out)), oracle.optic_frame(os_, spec, rad, rate, 10.0), "eager")
    snapshot = {k: getattr(ol, k).clone() for k in ("v", "adapt", "I_lp", "I_mean", "_fresh", "_cuda_dr")}
    pending = ol._pending_ms
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        ol.step_frame(rad_t, rate_t, 10.0, out=out)
    for k, v in snapshot.items():                 # capture records without running; keep the state as it was
        getattr(ol, k).copy_(v)
    ol._pending_ms = pending
    torch.cuda.synchronize()
    for rep in range(4):
        rad = rng.random((B, retina.n_col, 4)).astype(np.float32)
        rate = (rng.random((B, conn.n)) * 300.0).astype(np.float32)
        rad_t.copy_(T(rad))
        rate_t.copy_(T(rate))
        out.zero_()
        graph.replay()
        torch.cuda.synchronize()
        want = oracle.optic_frame(os_, spec, rad, rate, 10.0)
        assert_bits(N(out), want, f"drive replay {rep}")
        _check_state(ol, os_, spec, f"replay {rep}")
        torch.cuda.synchronize()
