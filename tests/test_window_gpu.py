"""The window stepper of the plain brain (csrc/window.hip, launch_lif_window): up to n_delay steps per launch from
64-step spike history words, against the CPU oracle's step-by-step reference sequence (decay, ascending-presynaptic
input of the spikes n_delay steps old, fused LIF update, flyverse/brain.py:830-870), bit for bit, with windows of
every length from 1 to n_delay that start at every phase of the delay ring, cells of every length class (short up to
one chunk, short in registers up to 1024 in-edges, long > 1024 on a workgroup, very long > 2048 and streamed > 6144
in-edges), Poisson forcing, counts and adaptation, and B = 1 or 3."""
import math

import numpy as np
import pandas as pd
import pytest
import scipy.sparse as sp
import torch

import oracle
from flyhip.brain import Brain, LIFParams, shaped_weights, window_length
from flyhip.connectome import Connectome

pytestmark = pytest.mark.gpu
DEV = "cuda"
CHUNKS = [1, 7, 18, 5, 30, 40, 18, 17, 2, 81]          # 219 steps; windows of every phase and length


def long_row_graph(seed=0, n=12_000, base_deg=40):
    """Signed integer-count graph with a Poisson(base_deg) in-degree plus rows of 700, 1,500, 3,000 and 11,000
    in-edges."""
    rng = np.random.default_rng(seed)
    deg = rng.poisson(base_deg, n)
    deg[10:15], deg[15:20], deg[20:25], deg[25] = 700, 1500, min(3000, n - 1), min(11_000, n - 1)
    post = np.repeat(np.arange(n), deg)
    pre = np.concatenate([rng.choice(n, d, replace=False) for d in deg])
    sign = np.where(rng.random(n) < 0.7, 1.0, -1.0)
    counts = rng.geometric(0.2, len(pre)).clip(1, 30)
    W = sp.csr_matrix((counts * sign[pre], (post, pre)), shape=(n, n), dtype=np.float32)
    W.sum_duplicates()
    neurons = pd.DataFrame({"bodyId": np.arange(n) + 10_000, "type": "", "superclass": "central"})
    return Connectome(neurons, W)


def N_(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("B,adapt", [(1, 0.0), (3, 0.0), (2, 1.5)])
def test_window_stepper_bit_exact_against_the_oracle(B, adapt):
    c = long_row_graph()
    lp = LIFParams.shiu(dt=0.1, adapt_jump=adapt)
    brain = Brain(c, lp, device=DEV, batch=B, seed=4321)
    assert brain._win_on and brain._win == window_length(brain.n_delay) == 18
    assert brain._n_long == 11 and brain._n_vlong == 6
    brain.record_activity = True
    rng = np.random.default_rng(1)
    inputs = np.sort(rng.choice(c.n, 400, replace=False))
    hz = rng.uniform(50, 400, 400).astype(np.float32)
    brain.set_poisson(inputs, hz)
    brain.freeze(np.arange(c.n - 100, c.n))                   # frozen cells that still receive input
    # the oracle: the same weights, the reference step sequence with the Philox uniforms of (seed, step, cell, fly)
    W = shaped_weights(c, lp)
    Wc = sp.csc_matrix(W)
    Wc.sum_duplicates()
    Wc.sort_indices()
    csc = (Wc.indptr.astype(np.int64), Wc.indices.astype(np.int32), Wc.data.astype(np.float32))
    st = oracle.BrainState(B, c.n, brain.n_delay, lp.v_rest)
    st.poisson_p[:] = N_(brain.poisson_p)
    st.active[:] = N_(brain.active)
    P = oracle.lif_params_vector(lp)
    a_s = math.exp(-lp.dt / lp.tau_syn)
    flags = 32 | (8 if adapt > 0 else 0)
    total = 0
    for chunk in CHUNKS:
        brain.step(chunk)
        for _ in range(chunk):
            oracle.brain_step(st, csc, P, a_s, 4321, flags)
            total += int(st.spikes.sum())
        torch.cuda.synchronize()
        for name in ("v", "g", "refrac", "rate", "spike_counts", "adapt", "spikes", "spike_buf"):
            got, want = N_(getattr(brain, name)), getattr(st, name)
            assert np.array_equal(got, want), (chunk, name, int((got != want).sum()))
        assert brain.buf_pos == st.buf_pos and brain.step_count == st.step_count
    assert total > 1000


def test_window_history_rebuilt_from_the_delay_ring():
    """After spike_buf is written from outside (a checkpoint restore), sync_bits rebuilds the history words, and a
    brain continued from a copied state matches the brain it was copied from."""
    c = long_row_graph(seed=2, n=3000)
    lp = LIFParams.shiu(dt=0.1)
    a = Brain(c, lp, device=DEV, batch=2, seed=9)
    b = Brain(c, lp, device=DEV, batch=2, seed=9)
    for br in (a, b):
        br.record_activity = True
        br.set_poisson(np.arange(0, 3000, 7), np.full(429, 300.0, np.float32))
    a.step(23)
    for name in ("v", "g", "refrac", "rate", "spike_counts", "spikes", "spike_buf", "drive", "adapt", "res"):
        getattr(b, name).copy_(getattr(a, name))
    b.buf_pos, b.step_count, b.t = a.buf_pos, a.step_count, a.t
    b.sync_bits()
    m = sum(1 << ((a.step_count - a.n_delay + j) % 64) for j in range(a.n_delay))    # the bits a window reads
    mask = torch.tensor(m - (1 << 64) if m >= 1 << 63 else m, dtype=torch.int64, device=DEV)
    assert torch.equal(a._hist & mask, b._hist & mask)
    a.step(50)
    b.step(50)
    for name in ("v", "g", "spike_counts", "spike_buf"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
