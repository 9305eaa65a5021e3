"""Pins the oracle's composite sequences (oracle.brain_step, oracle.optic_frame, oracle.FlyModel.frame) against the
reference's CPU step and frame written op by op with torch float32 CPU ops, the reference computation itself:
  * the LIF step of flyverse/brain.py:830-870 with the event-driven synaptic input of brain.py:679-692 (nonzero in
    (brain, pre) order, one index_add_), decay before the input;
  * OpticLobe._step_frame_inference (flyverse/optic.py:563-604) on the CPU branch: photoreceptor stage
    (optic.py:496-517), feedback from the brain's rates at frame start, the held products, `_substep` per substep
    (optic.py:525-551, dr recomputed from v each substep), delta_rate = rates() - r0, the clamped output drive;
  * the frame of FlyBrain._step_plain / _vision_frame (flyverse/fly.py:448-507): the frame's whole LIF steps and the
    carried fraction, the optic frame over the elapsed ms first, its drive COPIED into the brain, then the LIF steps.
The Poisson uniforms and the radiance movie are injected (the same Philox arrays both sides use); the sparse
products use the chunked order this build defines for them (checked here against torch's own sparse product).
Each frame compares spikes of every step and the whole state bit for bit."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from flyhip.brain import shaped_weights
from flyhip.optic import optic_matrices
from helpers import fly_setup, oracle_fly
from oracle.torch_ref import RefBrain, RefClockedBrain, RefOptic, chunked_mv, fly_frame_steps


def test_chunked_product_agrees_with_torch_sparse_product():
    """The defined order against the reference's own `_mv` (torch sparse CSR product, optic.py:161-165)."""
    rng = np.random.default_rng(3)
    M = sp.random(300, 500, density=0.08, random_state=rng, format="csr", dtype=np.float32)
    M.data -= 0.5
    x = torch.from_numpy(rng.normal(0, 1, (3, 500)).astype(np.float32))
    Mt = torch.sparse_csr_tensor(torch.from_numpy(M.indptr.astype(np.int64)),
                                 torch.from_numpy(M.indices.astype(np.int64)), torch.from_numpy(M.data), size=M.shape)
    torch.testing.assert_close(chunked_mv(M, x), (Mt @ x.T.contiguous()).T, rtol=1e-5, atol=1e-6)
    assert torch.equal(chunked_mv(M, x), torch.from_numpy(oracle.spmv(M.indptr, M.indices, M.data, x.numpy())))


@pytest.mark.parametrize("frames_ms", [[10.0] * 3, [7.35, 10.0, 2.5, 13.15, 0.05, 9.9]])
def test_fly_frames_oracle_equals_the_reference_torch_sequence(frames_ms):
    """oracle.FlyModel.frame (optic frame, drive copy, LIF steps) against the reference's frame sequence in torch ops,
    with frames whose lengths leave fractions of an LIF step and of an optic substep to carry."""
    conn, retina, layout, lp, op = fly_setup(0.01, seed=2)
    B, seeds, movie_seeds = 2, [5, 6], [7, 8]
    from flyhip import synth
    idx, hz = synth.fly_inputs(3, layout, 40, 400.0)
    om = oracle_fly(conn, retina, lp, op, B, seeds, movie_seeds)
    p_step = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    om.st.poisson_p[:, idx] = p_step
    # the reference side: the same host-built weights (shaped, rate units frozen and pruned, flyverse/fly.py:95-104)
    m = optic_matrices(conn, retina, op)
    keep = np.ones(conn.n, np.float32)
    keep[m["rate_idx"]] = 0.0
    W = (sp.diags(keep) @ shaped_weights(conn, lp) @ sp.diags(keep)).tocsr()
    W.eliminate_zeros()
    rb = RefBrain(W, B, lp, keep)
    rb.poisson_p[:, torch.from_numpy(idx)] = torch.from_numpy(hz) * (lp.dt / 1000.0)
    assert torch.equal(rb.poisson_p, torch.from_numpy(om.st.poisson_p))
    ro = RefOptic(m, op, retina, B, conn.n)
    pending = 0.0
    total_spikes = 0
    for frame, ms in enumerate(frames_ms):
        steps, pending = fly_frame_steps(pending, ms, lp.dt)
        if steps == 0:
            continue                                        # no LIF step: fly.py:465-466 returns before the optics
        movie_frame = om.frame_index
        raster = []
        om.frame(steps, raster)
        ncol = retina.n_col
        rad = oracle._per_fly_uniform(B, ncol * 4, movie_seeds, movie_frame, 1).reshape(B, ncol, 4)
        rb.drive.copy_(ro.step_frame(torch.from_numpy(rad), rb.rate, steps * lp.dt))
        assert torch.equal(rb.drive, torch.from_numpy(om.st.drive)), frame
        for k in range(steps):
            step = om.st.step_count - steps + k
            rb.step(torch.from_numpy(oracle._per_fly_uniform(B, conn.n, seeds, step, 0)))
            assert torch.equal(rb.spikes, torch.from_numpy(raster[k])), (frame, k)
            total_spikes += int(raster[k].sum())
        for name in ("v", "g", "refrac", "rate", "spike_counts", "adapt", "res"):
            assert torch.equal(getattr(rb, name), torch.from_numpy(getattr(om.st, name))), (frame, name)
        assert torch.equal(rb.spike_buf, torch.from_numpy(om.st.spike_buf)) and rb.buf_pos == om.st.buf_pos
        assert torch.equal(ro.v, torch.from_numpy(om.os.v)) and torch.equal(ro.adapt, torch.from_numpy(om.os.adapt))
        assert torch.equal(ro.I_lp, torch.from_numpy(om.os.I_lp))
        assert torch.equal(ro.I_mean, torch.from_numpy(om.os.I_mean))
        assert abs(ro._pending_ms - om.os.pending_ms) == 0.0
    assert total_spikes > 50                                    # the comparison saw real activity


@pytest.mark.parametrize("adapt,std", [(0.0, 0.0), (1.5, 0.0), (0.0, 0.3), (0.8, 0.2)])
def test_brain_step_oracle_equals_the_reference_torch_step(adapt, std):
    """oracle.brain_step (decay, event-driven input of the spikes n_delay steps old, injected Poisson uniforms, fused
    update, ring advance) against brain.py:830-870 in torch ops, with adaptation and short-term depression."""
    from flyhip import synth
    from flyhip.brain import LIFParams
    c, inputs, hz = synth.plain_lif_graph(4, n=600, n_inputs=60, hz=300.0)
    lp = LIFParams.shiu(dt=0.1, adapt_jump=adapt, std_u=std)
    W = shaped_weights(c, lp)
    B, seed = 3, 99
    st = oracle.BrainState(B, c.n, max(1, int(round(lp.delay / lp.dt))), lp.v_rest)
    st.std_u[:] = np.float32(std)
    st.poisson_p[:, inputs] = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    Wc = sp.csc_matrix(W)
    Wc.sum_duplicates()
    Wc.sort_indices()
    csc = (Wc.indptr.astype(np.int64), Wc.indices.astype(np.int32), Wc.data.astype(np.float32))
    rb = RefBrain(W, B, lp, np.ones(c.n, np.float32))
    rb.poisson_p.copy_(torch.from_numpy(st.poisson_p))
    flags = (8 if adapt > 0 else 0) | (2 if std > 0 else 0)
    P = oracle.lif_params_vector(lp)
    a_s = math.exp(-lp.dt / lp.tau_syn)
    fired = 0
    for step in range(250):
        rnd = torch.from_numpy(oracle.uniform(B, c.n, seed, step))
        oracle.brain_step(st, csc, P, a_s, seed, flags)
        rb.step(rnd)
        fired += int(st.spikes.sum())
        assert torch.equal(rb.spikes, torch.from_numpy(st.spikes)), step
        for name in ("v", "g", "refrac", "rate", "spike_counts", "adapt", "res"):
            assert torch.equal(getattr(rb, name), torch.from_numpy(getattr(st, name))), (step, name)
        assert torch.equal(rb.spike_buf, torch.from_numpy(st.spike_buf))
    assert fired > 100


@pytest.mark.parametrize("adapt", [0.0, 1.5])
def test_clocked_steps_oracle_equals_the_reference_torch_step(adapt):
    """The oracle's per-module clocks (FlyModel.set_clocks: phase coefficients, accumulators, the clocks' products
    and the fused update with flag 16) against Brain._step_clocked's torch branch (flyverse/brain.py:898-957) op by
    op, with clocks of 1, 2, 3 and 6 steps (period 6) and Poisson forcing."""
    from flyhip import synth
    from flyhip.brain import LIFParams
    c, inputs, hz = synth.plain_lif_graph(6, n=700, n_inputs=80, hz=400.0)
    lp = LIFParams.shiu(dt=0.1, adapt_jump=adapt)
    W = shaped_weights(c, lp)
    B, seed = 2, 31
    k = np.random.default_rng(2).choice([1, 2, 3, 6], c.n)
    om = oracle_fly(c, None, lp, None, B, seed, None, with_optic=False)
    om.flags = 32 | (8 if adapt > 0 else 0)
    om.st.poisson_p[:, inputs] = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    om.set_clocks(k, W, lp)
    rb = RefClockedBrain(W, B, lp, np.ones(c.n, np.float32))
    rb.set_clocks(k, W)
    rb.poisson_p.copy_(torch.from_numpy(om.st.poisson_p))
    fired = 0
    for step, st in enumerate(om.lif_steps(240)):
        rb.step(torch.from_numpy(oracle.uniform(B, c.n, seed, step)))
        fired += int(st.spikes.sum())
        for name in ("spikes", "v", "g", "refrac", "rate", "spike_counts", "adapt"):
            assert torch.equal(getattr(rb, name), torch.from_numpy(getattr(st, name))), (step, name)
        for kk, acc in om.clocks["acc"].items():
            assert torch.equal(rb._acc[kk], torch.from_numpy(acc)), (step, kk)
    assert fired > 100
