"""Builders shared by the tests: small seeded graphs and the oracle model of a flyhip FlyBrain."""
import math

import numpy as np
import scipy.sparse as sp

import oracle
from flyhip import synth
from flyhip.brain import LIFParams, shaped_weights
from flyhip.connectome import Retina
from flyhip.optic import OpticParams, optic_matrices


def random_csr(rng, rows, cols, density, long_rows=(), long_len=0, signed=True):
    M = sp.random(rows, cols, density=density, random_state=rng, format="csr", dtype=np.float32)
    if signed:
        M.data -= 0.5
    if long_rows:
        M = M.tolil()
        for r in long_rows:
            cs = rng.choice(cols, min(long_len, cols), replace=False)
            for c_ in cs:
                M[r, c_] = np.float32(rng.standard_normal())
        M = M.tocsr()
    M.sum_duplicates()
    M.sort_indices()
    return M


def fly_setup(scale=0.02, seed=1):
    conn, retina, layout = synth.fly_graph(seed, scale)
    lp = LIFParams.shiu(dt=0.1)
    op = OpticParams(tau_by_type=conn.rate_tau, baseline_by_type=conn.rate_baseline, pair_gain=[])
    return conn, retina, layout, lp, op


def oracle_fly(conn, retina, lp, op, B, seeds, movie_seeds, with_optic=True):
    """The CPU model of FlyBrain(conn, retina, lif_params=lp, optic_params=op) with the same host-built weights."""
    W = shaped_weights(conn, lp)
    active = np.ones(conn.n, np.float32)
    m = optic_matrices(conn, retina, op) if with_optic else None
    if m is not None:
        keep = np.ones(conn.n, np.float32)
        keep[m["rate_idx"]] = 0.0
        active[m["rate_idx"]] = 0.0
        W = (sp.diags(keep) @ W @ sp.diags(keep)).tocsr()
        W.eliminate_zeros()
    Wc = sp.csc_matrix(W)
    Wc.sum_duplicates()
    Wc.sort_indices()
    csc = (Wc.indptr.astype(np.int64), Wc.indices.astype(np.int32), Wc.data.astype(np.float32))
    P = oracle.lif_params_vector(lp)
    n_delay = max(1, int(round(lp.delay / lp.dt)))
    spec = oracle.OptSpec(m, op, retina.pr_sens, conn.n, retina.cells()) if m is not None else None
    model = oracle.FlyModel(conn.n, B, P, math.exp(-lp.dt / lp.tau_syn), n_delay, csc, active, seeds, spec,
                            movie_seeds, v_rest=lp.v_rest, dt=lp.dt)
    model.W = W                                   # the (post, pre) weights the model steps with
    return model


def grouped_retina(retina, per_column=3, seed=0):
    """The same photoreceptors placed `per_column` to a column in a shuffled order, with random spectral families and
    two columns left empty: cells that average several photoreceptors (flyverse/optic.py:320-325)."""
    rng = np.random.default_rng(seed)
    n = len(retina.pr_index)
    col = rng.permutation(n) // per_column
    return Retina(retina.pr_index, retina.pr_sens, col, rng.choice(5, n, p=[.6, .1, .1, .1, .1]), int(col.max()) + 3)


C4_TARGET_SEED, C4_READOUT_SEED = 9, 8


def c4_problem(T=None):
    """The full-size differentiable optic-lobe problem (89,390 units, ~6 M edges, 65 types, 32 grating movies,
    734 readout units, dt 5 ms): (W0, unit_type, P, target, readout, tau, bias, gain).  The network, movies and
    initial parameters are flyhip.optic_data's; the targets are seeded uniforms on [0, 0.3] (the readout rates' range)
    so that the problem is rebuilt identically without a GPU.  T: first T steps only (default all 200)."""
    from flyhip import optic_data as od
    T = od.T_STEPS if T is None else T
    W0, ut, W_rp, pos = od.optic_network()
    P = od.photoreceptor_drive(od.grating_intensity(pos, T), W_rp)
    readout = np.sort(np.random.default_rng(C4_READOUT_SEED).choice(W0.shape[0], od.N_READOUT, replace=False))
    target = np.random.default_rng(C4_TARGET_SEED).uniform(0.0, 0.3, (T, od.N_READOUT, od.MOVIES)).astype(np.float32)
    tau, bias, gain = od.initial_params()
    return W0, ut, P, target, readout, tau, bias, gain
